"""CPU oracle for the ORAC nonbonded force path -- TEST INFRASTRUCTURE ONLY.

Imported only by ``tests/``, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg, as the checker.  Two back ends:

* ``liboracle.so``: the plain-C FP64 restatement (``orac_oracle.c``), which
  also runs on the GPU host;
* ``_ref/libref.so``: the reference's own Fortran hot path, compiled in
  place by ``make ref`` in this container only, used to pin the restatement.

Every wrapper takes arrays in the reference's layouts (see ``orac_oracle.h``).
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        _lib = C.CDLL(path)
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref.so"))


def ref():
    """The compiled reference, or None when it was not built here."""
    global _ref
    if _ref is None and ref_available():
        _ref = C.CDLL(os.path.join(HERE, "_ref", "libref.so"))
        _ref.refglue_init()
    return _ref


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def _i(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel(order="F"))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def recip_of(oc):
    """recip(i,j) = 0.5*oc(j,i) (fft_pme.f:47-51), Fortran order."""
    oc = np.asarray(oc, dtype=np.float64)
    return np.asfortranarray(0.5 * oc.T)


# ---------------------------------------------------------------- oracle

def bsp_moduli(order, nfft):
    out = np.zeros(nfft)
    lib().orc_load_bsp_moduli(C.c_int(order), C.c_int(nfft), _p(out))
    return out


def fill_bspline(w, order):
    a, d = np.zeros(order), np.zeros(order)
    lib().orc_fill_bspline(C.c_double(w), C.c_int(order), _p(a), _p(d))
    return a, d


def charge_grid(x, y, z, q, recip, nfft, order, nthreads=1):
    """Fractionals, B-splines and the spread grid Q (K3,K2,K1 C-order view)."""
    n = len(x)
    k1, k2, k3 = nfft
    x, y, z, q, r = _f(x), _f(y), _f(z), _f(q), _f(recip)
    fr = [np.zeros(n) for _ in range(3)]
    L = lib()
    L.orc_scaled_fractionals(C.c_int(n), _p(x), _p(y), _p(z), _p(r), C.c_int(k1), C.c_int(k2), C.c_int(k3),
                             *[_p(a) for a in fr])
    th = [np.zeros(n * order) for _ in range(6)]
    L.orc_bspline_coeffs(C.c_int(n), *[_p(a) for a in fr], C.c_int(order), *[_p(a) for a in th])
    Q = np.zeros(k1 * k2 * k3)
    L.orc_fill_charge_grid(C.c_int(n), _p(q), _p(th[0]), _p(th[1]), _p(th[2]), *[_p(a) for a in fr],
                           C.c_int(order), C.c_int(k1), C.c_int(k2), C.c_int(k3), _p(Q), C.c_int(nthreads))
    return dict(fr=fr, theta=th[:3], dtheta=th[3:], Q=Q.reshape(k3, k2, k1))


def fft_r2c(Q, nthreads=1):
    k3, k2, k1 = Q.shape
    FQ = np.zeros((k3, k2, k1 // 2 + 1), dtype=np.complex128)
    Qc = np.ascontiguousarray(Q, dtype=np.float64)
    lib().orc_fft_r2c(C.c_int(k1), C.c_int(k2), C.c_int(k3), _p(Qc), _p(FQ), C.c_int(nthreads))
    return FQ


def fft_c2r(FQ, k1, nthreads=1):
    k3, k2, _ = FQ.shape
    Q = np.zeros((k3, k2, k1))
    F = np.ascontiguousarray(FQ, dtype=np.complex128)
    lib().orc_fft_c2r(C.c_int(k1), C.c_int(k2), C.c_int(k3), _p(F), _p(Q), C.c_int(nthreads))
    return Q


def scalar_sum(FQ, alpha, volume, recip, bsp, nfft, rkcut=1e9, nthreads=1):
    """FFTW-variant scalar_sum; returns (eer, vir(3,3), FQ scaled)."""
    F = np.ascontiguousarray(FQ, dtype=np.complex128).copy()
    vir = np.zeros(9)
    eer = C.c_double()
    lib().orc_scalar_sum(_p(F), C.c_double(alpha), C.c_double(volume), _p(_f(recip)), _p(_f(bsp[0])),
                         _p(_f(bsp[1])), _p(_f(bsp[2])), C.c_int(nfft[0]), C.c_int(nfft[1]), C.c_int(nfft[2]),
                         C.c_double(rkcut), C.byref(eer), _p(vir), C.c_int(nthreads))
    return eer.value, vir.reshape(3, 3, order="F"), F


def fft_pme(s, nthreads=1, pressure=True, rkcut=1e9, fp=None, stressc=None, variant="fftw"):
    """fft_pme.f semantics on System s: returns eer, forces (+= fp), stressc (+=).

    variant 'fftw' is the production FFTW branch; 'fftpack' the reference's
    fftpack branch (full complex grid), which the compiled reference runs.
    """
    n = s.natoms
    cm = s.molecule_coms()
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)] if fp is None else [_f(a).copy() for a in fp]
    st = np.zeros(9) if stressc is None else _f(stressc).copy()
    bsp = [bsp_moduli(s.order, k) for k in s.nfft]
    eer = C.c_double()
    lib().orc_fft_pme(C.c_int(n), _p(_f(s.xp0)), _p(_f(s.yp0)), _p(_f(s.zp0)), _p(_f(cm[0])), _p(_f(cm[1])),
                      _p(_f(cm[2])), _p(_f(s.charge)), _p(_f(s.co)), _p(_f(s.oc)), C.c_double(s.volume),
                      C.c_double(s.alpha), C.c_int(s.order), C.c_int(s.nfft[0]), C.c_int(s.nfft[1]),
                      C.c_int(s.nfft[2]), C.byref(eer), _p(fpx), _p(fpy), _p(fpz), _p(st), _p(_i(s.atomp)),
                      C.c_int32(1 if pressure else 0), C.c_double(rkcut), _p(bsp[0]), _p(bsp[1]), _p(bsp[2]),
                      C.c_int(nthreads), C.c_int(0 if variant == "fftw" else 1))
    return dict(eer=eer.value, f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"))


def last_pme_times():
    t = np.zeros(5)
    lib().orc_last_pme_times(_p(t))
    return t


def lc_index(ncx, ncy, ncz, rcut, co, indmax=2500):
    nind = np.zeros(2, dtype=np.int32)
    ii, jj, kk = [np.zeros(2 * indmax, dtype=np.int32) for _ in range(3)]
    rc = lib().orc_lc_index(C.c_int(indmax), C.c_int(ncx), C.c_int(ncy), C.c_int(ncz), _p(nind), _p(ii), _p(jj),
                            _p(kk), C.c_double(rcut), _p(_f(co)))
    if rc != 0:
        raise ValueError("lc_index: stencil overflow or a cell would be visited twice")
    return nind, ii, jj, kk


def lc_list(ncx, ncy, ncz, stencil, rcut, co, xpg, ypg, zpg, capacity=None):
    nind, ii, jj, kk = stencil
    ng = len(xpg)
    cap = capacity or (ng * 64 + 16)
    while True:
        out = np.zeros(cap, dtype=np.int32)
        total = lib().orc_lc_list(C.c_int(ncx), C.c_int(ncy), C.c_int(ncz), _p(nind), _p(ii), _p(jj), _p(kk),
                                  C.c_double(rcut), _p(_f(co)), _p(_f(xpg)), _p(_f(ypg)), _p(_f(zpg)),
                                  C.c_int(ng), _p(out), C.c_long(cap))
        if total >= 0:
            return out[:total]
        cap *= 2


def all_groups_list(ngrp):
    """[ngrp, 1..ngrp]: the list the O(N^2) 'u' update call walks (iz = 0)."""
    return np.concatenate([[ngrp], np.arange(1, ngrp + 1)]).astype(np.int32)


class ForceArgs:
    """The arrays force.f consumes for System s, in Fortran layouts."""

    def __init__(self, s):
        self.s = s
        xpg = s.group_centres()
        cm = s.molecule_coms()
        self.xpg, self.ypg, self.zpg = [_f(a) for a in xpg]
        self.xpcm, self.ypcm, self.zpcm = [_f(a) for a in cm]
        self.mapnl = _i(s.mapnl())
        self.tt = _i(s.type_table)
        self.grppt = _i(s.grppt)


def force(s, nnlpp, *, rinn, rout, rtolinn, rtolout, rneigh, iz=1, lskip_ewald=False, ewald=True,
          nthreads=1, fa=None, massinfty=1.0e20, fp=None, stress=None, nnlpp1_cap=None):
    fa = fa or ForceArgs(s)
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)] if fp is None else [_f(a).copy() for a in fp]
    st = np.zeros(9) if stress is None else _f(stress).copy()
    e = [C.c_double() for _ in range(6)]
    cap = nnlpp1_cap or (len(nnlpp) + s.ngrp + 16 if iz == 1 else s.ngrp * (s.ngrp // 2 + 2))
    nn1 = np.zeros(cap, dtype=np.int32)
    nc1 = C.c_int32()
    rti = C.c_double(rtolinn)
    rc = lib().orc_force(
        _p(_i(s.ss_index)), _p(_f(s.xp0)), _p(_f(s.yp0)), _p(_f(s.zp0)), _p(fa.xpg), _p(fa.ypg), _p(fa.zpg),
        _p(_f(s.charge)), _p(_i(s.nbtype)), _p(fa.tt), C.c_int(s.ntypes), C.c_int(n), _p(_i(s.atomg)),
        _p(fa.xpcm), _p(fa.ypcm), _p(fa.zpcm), _p(_i(s.atomp)), _p(_f(s.co)), _p(_f(s.ecc12)), _p(_f(s.ecc6)),
        C.c_double(s.alpha), _p(fa.mapnl), C.c_int(s.ngrp), _p(fa.grppt), *[C.byref(x) for x in e],
        _p(fpx), _p(fpy), _p(fpz), _p(st), _p(_i(nnlpp)), _p(nn1), C.c_long(cap), C.c_double(rneigh),
        C.c_double(rinn), C.c_double(rout), C.byref(rti), C.c_double(rtolout), _p(_f(s.pmass)), C.c_int(iz),
        C.byref(nc1), _p(_f(s.mass)), C.c_double(massinfty), C.c_int32(int(lskip_ewald)), C.c_int32(int(ewald)),
        C.c_int(nthreads))
    if rc != 0:
        raise RuntimeError("orc_force: nnlpp1 capacity exceeded")
    names = ["uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"]
    out = {k: v.value for k, v in zip(names, e)}
    out.update(f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"), nnlpp1=nn1[:nc1.value].copy(),
               ncount1=nc1.value, rtolinn=rti.value)
    return out


def force_rem(s, nnlpp, rematom, *, rem_kind, rem_f3, rinn, rout, rtolinn, rtolout, rneigh, iz=1,
              lskip_ewald=False, ewald=True, lskip_force=False, ljintra=False, fa=None, massinfty=1.0e20):
    """force_REM.f on System s: force() plus the segment scaling of the pairs
    seg_select(rem_kind, rematom) picks; returns force()'s keys and seg_ene."""
    fa = fa or ForceArgs(s)
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    st = np.zeros(9)
    e = [C.c_double() for _ in range(6)]
    cap = len(nnlpp) + s.ngrp + 16 if iz == 1 else s.ngrp * (s.ngrp // 2 + 2)
    nn1 = np.zeros(cap, dtype=np.int32)
    nc1 = C.c_int32()
    rti = C.c_double(rtolinn)
    seg = C.c_double()
    rc = lib().orc_force_rem(
        _p(_i(s.ss_index)), _p(_f(s.xp0)), _p(_f(s.yp0)), _p(_f(s.zp0)), _p(fa.xpg), _p(fa.ypg), _p(fa.zpg),
        _p(_f(s.charge)), _p(_i(s.nbtype)), _p(fa.tt), C.c_int(s.ntypes), C.c_int(n), _p(_i(s.atomg)),
        _p(fa.xpcm), _p(fa.ypcm), _p(fa.zpcm), _p(_i(s.atomp)), _p(_f(s.co)), _p(_f(s.ecc12)), _p(_f(s.ecc6)),
        C.c_double(s.alpha), _p(fa.mapnl), C.c_int(s.ngrp), _p(fa.grppt), *[C.byref(x) for x in e],
        _p(fpx), _p(fpy), _p(fpz), _p(st), _p(_i(nnlpp)), _p(nn1), C.c_long(cap), C.c_double(rneigh),
        C.c_double(rinn), C.c_double(rout), C.byref(rti), C.c_double(rtolout), _p(_f(s.pmass)), C.c_int(iz),
        C.byref(nc1), _p(_f(s.mass)), C.c_double(massinfty), C.byref(seg), _p(_i(rematom)), C.c_int(rem_kind),
        C.c_double(rem_f3), C.c_int32(int(lskip_force)), C.c_int32(int(ljintra)), C.c_int32(int(lskip_ewald)),
        C.c_int32(int(ewald)))
    if rc != 0:
        raise RuntimeError("orc_force_rem: nnlpp1 capacity exceeded")
    names = ["uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"]
    out = {k: v.value for k, v in zip(names, e)}
    out.update(f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"), nnlpp1=nn1[:nc1.value].copy(),
               ncount1=nc1.value, rtolinn=rti.value, seg_ene=seg.value)
    return out


def decide_force(rinn, alpha):
    return bool(lib().orc_decide_force(C.c_double(rinn), C.c_double(alpha)))


def ferrf(s, pairs, iz, fudge=1.0, fp=None, erf=None):
    """ferrf.f; erf = (erf_arr_corr (4, n), delew, rlew) selects the ERF_CORR branch."""
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)] if fp is None else [_f(a).copy() for a in fp]
    a, b = C.c_double(), C.c_double()
    lst = _i(pairs)
    tab = _f(erf[0]) if erf is not None else np.zeros(4)
    rc = lib().orc_ferrf(_p(_i(s.ss_index)), C.c_double(s.alpha), _p(_f(s.charge)), C.c_double(fudge), _p(_f(x)),
                         _p(_f(y)), _p(_f(z)), C.c_int(iz), _p(lst), C.c_int(lst.size // 2), C.byref(a), C.byref(b),
                         _p(fpx), _p(fpy), _p(fpz), C.c_int32(0 if erf is None else 1), C.c_int(0), _p(tab),
                         C.c_double(erf[1] if erf is not None else 0.0), C.c_double(erf[2] if erf is not None else 0.0))
    assert rc == 0
    return dict(slt=a.value, slv=b.value, f=np.stack([fpx, fpy, fpz]))


def fnbgrp(s, ingrp, typ, ewald):
    """fnbgrp.f on the Cartesian coordinates of System s: pairs ingrp (2, n)
    1-based, per-atom LJ types typ (1-based, packed j(j-1)/2 + i tables)."""
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    lst = _i(ingrp)
    lib().orc_fnbgrp(_p(_i(s.ss_index)), _p(_f(x)), _p(_f(y)), _p(_f(z)), _p(_f(s.charge)), _p(_i(typ)),
                     _p(_f(s.ecc12)), _p(_f(s.ecc6)), C.c_int32(int(ewald)), C.c_double(s.alpha), _p(lst),
                     C.c_int(lst.size // 2), _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e])
    return dict(uconf_slt=e[0].value, uconf_slv=e[1].value, ucoul_slt=e[2].value, ucoul_slv=e[3].value,
                f=np.stack([fpx, fpy, fpz]))


def cself(s, kcut=1e9):
    a, b = C.c_double(), C.c_double()
    lib().orc_cself(_p(_i(s.ss_index)), C.c_int(s.natoms), C.c_double(s.alpha), C.c_double(kcut), _p(_f(s.charge)),
                    C.byref(a), C.byref(b))
    return a.value, b.value


# ---------------------------------------------------------------- reference

def _big_stack(fn, *args):
    """Run fn on a thread with a 1 GiB stack (the reference's fft_pme.f keeps
    ~75 MB of fixed-size local arrays)."""
    res = {}

    def run():
        try:
            res["v"] = fn(*args)
        except BaseException as e:  # pragma: no cover
            res["e"] = e

    old = threading.stack_size(1 << 30)
    try:
        t = threading.Thread(target=run)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in res:
        raise res["e"]
    return res.get("v")


def _ri(v):
    return C.byref(C.c_int32(int(v)))


def _rd(v):
    return C.byref(C.c_double(float(v)))


def ref_bsp_moduli(order, nfft, natoms=1):
    R = ref()
    R.refglue_pme_init(C.c_int(natoms), C.c_int(nfft[0]), C.c_int(nfft[1]), C.c_int(nfft[2]), C.c_int(order))
    b = [np.zeros(k) for k in nfft]
    R.refglue_get_bsp(_p(b[0]), _p(b[1]), _p(b[2]), C.c_int(nfft[0]), C.c_int(nfft[1]), C.c_int(nfft[2]))
    return b


def ref_fft_pme(s, pressure=True, rkcut=1e9):
    """The reference fft_pme (fftpack branch of do_pmesh_kspace)."""
    R = ref()
    n = s.natoms
    R.refglue_pme_init(C.c_int(n), C.c_int(s.nfft[0]), C.c_int(s.nfft[1]), C.c_int(s.nfft[2]), C.c_int(s.order))
    cm = s.molecule_coms()
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    st = np.zeros(9)
    eer = C.c_double()
    arrs = [_f(s.xp0), _f(s.yp0), _f(s.zp0), _f(cm[0]), _f(cm[1]), _f(cm[2]), _f(s.charge), _f(s.co), _f(s.oc)]
    atomp, grppt = _i(s.atomp), _i(s.grppt)

    def call():
        R.fft_pme_(_ri(n), *[_p(a) for a in arrs[:7]], _p(arrs[7]), _p(arrs[8]), _rd(s.volume), _rd(s.alpha),
                   _ri(s.order), _ri(s.nfft[0]), _ri(s.nfft[1]), _ri(s.nfft[2]), C.byref(eer), _p(fpx), _p(fpy),
                   _p(fpz), _p(st), _p(atomp), _p(grppt), _ri(1 if pressure else 0), _rd(rkcut))

    _big_stack(call)
    return dict(eer=eer.value, f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"))


def ref_scalar_sum_fftw(FQ, alpha, volume, recip, bsp, nfft, rkcut=1e9):
    R = ref()
    F = np.ascontiguousarray(FQ, dtype=np.complex128).copy()
    vir = np.zeros(9)
    eer = C.c_double()
    R.scalar_sum_fftw_(_p(F), _rd(alpha), _rd(volume), _p(_f(recip)), _p(_f(bsp[0])), _p(_f(bsp[1])),
                       _p(_f(bsp[2])), _ri(nfft[0]), _ri(nfft[1]), _ri(nfft[2]), C.byref(eer), _p(vir), _rd(rkcut))
    return eer.value, vir.reshape(3, 3, order="F"), F


def ref_lc_index(ncx, ncy, ncz, rcut, co, indmax=2500):
    R = ref()
    nind = np.zeros(2, dtype=np.int32)
    ii, jj, kk = [np.zeros(2 * indmax, dtype=np.int32) for _ in range(3)]
    R.lc_index_(_ri(indmax), _ri(ncx), _ri(ncy), _ri(ncz), _p(nind), _p(ii), _p(jj), _p(kk), _rd(rcut), _p(_f(co)))
    return nind, ii, jj, kk


def ref_lc_list(ncx, ncy, ncz, stencil, rcut, co, xpg, ypg, zpg):
    R = ref()
    nind, ii, jj, kk = stencil
    ng = len(xpg)
    out = np.zeros(ng * 400 + 16, dtype=np.int32)
    R.lc_list_(_ri(ncx), _ri(ncy), _ri(ncz), _p(nind), _p(ii), _p(jj), _p(kk), _rd(rcut), _p(_f(co)), _p(_f(xpg)),
               _p(_f(ypg)), _p(_f(zpg)), _ri(ng), _p(out), _ri(6))
    # walk the per-group [nv, l...] layout to find its length
    pos = 0
    for _ in range(ng):
        pos += int(out[pos]) + 1
    return out[:pos].copy()


def ref_force(s, nnlpp, *, rinn, rout, rtolinn, rtolout, rneigh, iz=1, lskip_ewald=False, ewald=True,
              massinfty=1.0e20, fa=None):
    R = ref()
    fa = fa or ForceArgs(s)
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    st = np.zeros(9)
    e = [C.c_double() for _ in range(6)]
    cap = len(nnlpp) + s.ngrp + 16 if iz == 1 else s.ngrp * (s.ngrp // 2 + 2)
    nn1 = np.zeros(cap, dtype=np.int32)
    nc1 = C.c_int32()
    rti = C.c_double(rtolinn)
    keep = [_i(s.ss_index), _f(s.xp0), _f(s.yp0), _f(s.zp0), _f(s.charge), _i(s.nbtype), _i(s.atomg),
            _i(s.atomp), _f(s.co), _f(s.ecc12), _f(s.ecc6), _i(nnlpp), _f(s.pmass), _f(s.mass),
            np.zeros(s.ngrp, dtype=np.int32)]
    (ss, x, y, z, q, nbt, atg, atp, co, e12, e6, nl, pm, ms, groupp) = keep

    def call():
        R.force_(_p(ss), _p(x), _p(y), _p(z), _p(fa.xpg), _p(fa.ypg), _p(fa.zpg), _p(q), _p(nbt), _p(fa.tt),
                 _ri(s.ntypes), _ri(n), _p(atg), _p(fa.xpcm), _p(fa.ypcm), _p(fa.zpcm), _p(groupp), _p(atp),
                 _p(co), _p(e12), _p(e6), _rd(s.alpha), _p(fa.mapnl), _ri(s.ngrp), _p(fa.grppt),
                 *[C.byref(v) for v in e], _p(fpx), _p(fpy), _p(fpz), _p(st), _p(nl), _p(nn1), _rd(rneigh),
                 _rd(rinn), _rd(rout), C.byref(rti), _rd(rtolout), _p(pm), _ri(iz), C.byref(nc1), _p(ms),
                 _rd(massinfty), _ri(int(lskip_ewald)), _ri(int(ewald)))

    _big_stack(call)
    names = ["uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"]
    out = {k: v.value for k, v in zip(names, e)}
    out.update(f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"), nnlpp1=nn1[:nc1.value].copy(),
               ncount1=nc1.value, rtolinn=rti.value)
    return out


def ref_force_rem(s, nnlpp, rematom, *, rem_kind, rem_f3, rinn, rout, rtolinn, rtolout, rneigh, iz=1,
                  lskip_ewald=False, ewald=True, lskip_force=False, ljintra=False, massinfty=1.0e20, fa=None):
    """The reference's force_rem (force_REM.f, serial) with module rem's
    rem_segkind, rem_factor(3) and lskip_force, and cpropar's ljintra set."""
    R = ref()
    R.refglue_set_rem(C.c_int(int(rem_kind)), C.c_double(float(rem_f3)), C.c_int(int(lskip_force)),
                      C.c_int(int(ljintra)))
    fa = fa or ForceArgs(s)
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    st = np.zeros(9)
    e = [C.c_double() for _ in range(6)]
    cap = len(nnlpp) + s.ngrp + 16 if iz == 1 else s.ngrp * (s.ngrp // 2 + 2)
    nn1 = np.zeros(cap, dtype=np.int32)
    nc1 = C.c_int32()
    rti = C.c_double(rtolinn)
    seg = C.c_double()
    keep = [_i(s.ss_index), _f(s.xp0), _f(s.yp0), _f(s.zp0), _f(s.charge).copy(), _i(s.nbtype), _i(s.atomg),
            _i(s.atomp), _f(s.co), _f(s.ecc12), _f(s.ecc6), _i(nnlpp), _f(s.pmass), _f(s.mass),
            np.zeros(s.ngrp, dtype=np.int32), _i(rematom)]
    (ss, x, y, z, q, nbt, atg, atp, co, e12, e6, nl, pm, ms, groupp, rem) = keep

    def call():
        R.force_rem_(_p(ss), _p(x), _p(y), _p(z), _p(fa.xpg), _p(fa.ypg), _p(fa.zpg), _p(q), _p(nbt), _p(fa.tt),
                     _ri(s.ntypes), _ri(n), _p(atg), _p(fa.xpcm), _p(fa.ypcm), _p(fa.zpcm), _p(groupp), _p(atp),
                     _p(co), _p(e12), _p(e6), _rd(s.alpha), _p(fa.mapnl), _ri(s.ngrp), _p(fa.grppt),
                     *[C.byref(v) for v in e], _p(fpx), _p(fpy), _p(fpz), _p(st), _p(nl), _p(nn1), _rd(rneigh),
                     _rd(rinn), _rd(rout), C.byref(rti), _rd(rtolout), _p(pm), _ri(iz), C.byref(nc1), _p(ms),
                     _rd(massinfty), C.byref(seg), _p(rem), _ri(int(lskip_ewald)), _ri(int(ewald)))

    _big_stack(call)
    names = ["uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"]
    out = {k: v.value for k, v in zip(names, e)}
    out.update(f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"), nnlpp1=nn1[:nc1.value].copy(),
               ncount1=nc1.value, rtolinn=rti.value, seg_ene=seg.value)
    return out


def ref_ferrf(s, pairs, iz, fudge=1.0, erf=None):
    R = ref()
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    lst = _i(pairs)
    tab = _f(erf[0]) if erf is not None else np.zeros(8)
    keep = [_i(s.ss_index), _f(s.charge), _f(x), _f(y), _f(z)]
    R.ferrf_(_p(keep[0]), _rd(s.alpha), _p(keep[1]), _rd(fudge), _p(keep[2]), _p(keep[3]), _p(keep[4]), _ri(iz),
             _p(lst), _ri(lst.size // 2), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz),
             _ri(0 if erf is None else 1), _p(tab), _rd(erf[1] if erf is not None else 0.0),
             _rd(erf[2] if erf is not None else 0.0), _ri(0), _ri(n))
    return dict(slt=a.value, slv=b.value, f=np.stack([fpx, fpy, fpz]))


def ref_fnbgrp(s, ingrp, typ, ewald):
    """The reference fnbgrp (serial) with the arguments of fnbgrp() above."""
    R = ref()
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    keep = [_i(s.ss_index), _f(x), _f(y), _f(z), _f(s.charge), _i(typ), _f(s.ecc12), _f(s.ecc6), _i(ingrp)]
    R.fnbgrp_(*[_p(k) for k in keep[:8]], _ri(int(ewald)), _rd(s.alpha), _p(keep[8]), _ri(keep[8].size // 2),
              _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e], _ri(n), _ri(1))
    return dict(uconf_slt=e[0].value, uconf_slv=e[1].value, ucoul_slt=e[2].value, ucoul_slv=e[3].value,
                f=np.stack([fpx, fpy, fpz]))


def ref_cself(s, kcut=1e9):
    R = ref()
    a, b = C.c_double(), C.c_double()
    R.cself_(_p(_i(s.ss_index)), _ri(s.natoms), _rd(s.alpha), _rd(kcut), _p(_f(s.charge)), C.byref(a), C.byref(b))
    return a.value, b.value


# ---------------------------------------------------------------- alchemy

E6 = ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss")


def alch_set(ladd, lrm, intramol=False, ljintra=False, qq=False):
    """alch_fact(1:2, -4:4) as an (9, 2) array indexed [typeaij + 4]."""
    f = np.zeros(18)
    lib().orc_alch_set(_p(f), C.c_double(ladd), C.c_double(lrm), C.c_int(int(intramol)), C.c_int(int(ljintra)),
                       C.c_int(int(qq)))
    return f.reshape(9, 2)


def force_alchemy(s, al, lam, nnlpp, *, rinn, rout, rtolinn, rtolout, rneigh, iz=1, lskip_ewald=False,
                  ewald=True, fa=None, massinfty=1.0e20, dwrk=0.0):
    """force_alchemy.f (serial) on System s with Alchemy al and the eight
    lambdas lam (synth.lambda_state order)."""
    fa = fa or ForceArgs(s)
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    st = np.zeros(9)
    e6 = np.zeros(6)
    cap = len(nnlpp) + s.ngrp + 16 if iz == 1 else s.ngrp * (s.ngrp // 2 + 2)
    nn1 = np.zeros(cap, dtype=np.int32)
    nc1 = C.c_int32()
    rti = C.c_double(rtolinn)
    dw = C.c_double(dwrk)
    rc = lib().orc_force_alchemy(
        _p(_i(s.ss_index)), _p(_f(s.xp0)), _p(_f(s.yp0)), _p(_f(s.zp0)), _p(fa.xpg), _p(fa.ypg), _p(fa.zpg),
        _p(_f(s.charge)), _p(_i(al.ss_alchemy)), _p(_f(al.rmin)), _p(_f(al.epsm)), C.c_int(n), _p(fa.xpcm),
        _p(fa.ypcm), _p(fa.zpcm), _p(_i(s.atomp)), _p(_f(s.co)), C.c_double(s.alpha), _p(fa.mapnl),
        C.c_int(s.ngrp), _p(fa.grppt), _p(e6), _p(fpx), _p(fpy), _p(fpz), _p(st), _p(_i(nnlpp)), _p(nn1),
        C.c_long(cap), C.c_double(rneigh), C.c_double(rinn), C.c_double(rout), C.byref(rti), C.c_double(rtolout),
        _p(_f(s.pmass)), C.c_int(iz), C.byref(nc1), _p(_f(s.mass)), C.c_double(massinfty), _p(_f(lam)),
        C.c_int(int(al.intramol)), C.c_int(int(al.ljintra)), C.c_int(int(al.opls)), C.byref(dw),
        C.c_int32(int(lskip_ewald)), C.c_int32(int(ewald)), C.c_double(al.soft_params()[0]),
        C.c_double(al.soft_params()[1]), C.c_int(int(al.shifted)))
    if rc != 0:
        raise RuntimeError("orc_force_alchemy: nnlpp1 capacity exceeded")
    out = dict(zip(E6, e6.tolist()))
    out.update(f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"), dwrk=dw.value,
               nnlpp1=nn1[:nc1.value].copy(), ncount1=nc1.value)
    return out


def ferrf_alchemy(s, charge, charge0, pairs, iz, fudge=1.0):
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, dw = C.c_double(), C.c_double(), C.c_double()
    lst = _i(pairs)
    lib().orc_ferrf_alchemy(_p(_i(s.ss_index)), C.c_double(s.alpha), _p(_f(charge)), _p(_f(charge0)),
                            C.c_double(fudge), _p(_f(x)), _p(_f(y)), _p(_f(z)), C.c_int(iz), _p(lst),
                            C.c_int(lst.size // 2), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), C.c_int(0),
                            C.byref(dw))
    return dict(slt=a.value, slv=b.value, dwrk=dw.value, f=np.stack([fpx, fpy, fpz]))


def cself_alchemy(s, charge, charge0, kcut=1e9):
    a, dw = C.c_double(), C.c_double()
    lib().orc_cself_alchemy(_p(_i(s.ss_index)), C.c_int(s.natoms), C.c_double(s.alpha), C.c_double(kcut),
                            _p(_f(charge)), _p(_f(charge0)), C.byref(a), C.byref(dw))
    return a.value, dw.value


def fft_pme_alchemy(s, charge, charge0, variant="fftw", nthreads=1, rkcut=1e9):
    """do_pmesh_alchemy.f90: the reciprocal sum on `charge` (energy, forces,
    virial) and the energy eer0 on `charge0` from the same B-spline weights and
    influence function (scalar_sum_alchemy.f90: the same terms summed in the
    same order as scalar_sum); dwrk = eer - eer0."""
    import dataclasses
    r = fft_pme(dataclasses.replace(s, charge=np.asarray(charge, float)), nthreads=nthreads, rkcut=rkcut,
                variant=variant)
    r0 = fft_pme(dataclasses.replace(s, charge=np.asarray(charge0, float)), nthreads=nthreads, rkcut=rkcut,
                 variant=variant)
    r.update(eer0=r0["eer"], dwrk=r["eer"] - r0["eer"])
    return r


def ref_set_alchemy(al, lam):
    ref().refglue_set_alchemy(C.c_int(len(al.ss_alchemy)), _p(_i(al.ss_alchemy)), _p(_f(lam)),
                              C.c_int(int(al.intramol)), C.c_int(int(al.ljintra)), C.c_int(int(al.opls)))


def ref_force_alchemy(s, al, lam, nnlpp, *, rinn, rout, rtolinn, rtolout, rneigh, iz=1, lskip_ewald=False,
                      ewald=True, massinfty=1.0e20, fa=None, dwrk=0.0):
    """The reference force_alchemy (serial build) with the arguments of force_alchemy() above."""
    R = ref()
    ref_set_alchemy(al, lam)
    fa = fa or ForceArgs(s)
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    st = np.zeros(9)
    e = [C.c_double() for _ in range(6)]
    cap = len(nnlpp) + s.ngrp + 16 if iz == 1 else s.ngrp * (s.ngrp // 2 + 2)
    nn1 = np.zeros(cap, dtype=np.int32)
    nc1 = C.c_int32()
    rti = C.c_double(rtolinn)
    dw = C.c_double(dwrk)
    lam_dummy = np.zeros(n)
    keep = [_i(s.ss_index), _f(s.xp0), _f(s.yp0), _f(s.zp0), _f(s.charge), _i(s.nbtype), _i(s.atomg),
            _i(s.atomp), _f(s.co), _f(s.ecc12), _f(s.ecc6), _i(nnlpp), _f(s.pmass), _f(s.mass),
            np.zeros(s.ngrp, dtype=np.int32), _f(al.rmin), _f(al.epsm)]
    (ss, x, y, z, q, nbt, atg, atp, co, e12, e6, nl, pm, ms, groupp, rmin, epsm) = keep

    fn = R.force_alchemy_shift_ if al.shifted else R.force_alchemy_
    qx, ljx = al.soft_params()

    def call():
        fn(_p(ss), _p(x), _p(y), _p(z), _p(fa.xpg), _p(fa.ypg), _p(fa.zpg), _p(q), _p(nbt),
                         _p(fa.tt), _ri(s.ntypes), _ri(n), _p(atg), _p(fa.xpcm), _p(fa.ypcm), _p(fa.zpcm),
                         _p(groupp), _p(atp), _p(co), _p(e12), _p(e6), _rd(s.alpha), _p(fa.mapnl), _ri(s.ngrp),
                         _p(fa.grppt), *[C.byref(v) for v in e], _p(fpx), _p(fpy), _p(fpz), _p(st), _p(nl),
                         _p(nn1), _rd(rneigh), _rd(rinn), _rd(rout), C.byref(rti), _rd(rtolout), _p(pm), _ri(iz),
                         C.byref(nc1), _p(ms), _rd(massinfty), _p(lam_dummy), _p(lam_dummy), _p(lam_dummy),
                         _p(lam_dummy), _p(rmin), _p(epsm), C.byref(dw), _ri(int(lskip_ewald)), _ri(int(ewald)),
                         _rd(qx), _rd(ljx))

    _big_stack(call)
    out = {k: v.value for k, v in zip(E6, e)}
    out.update(f=np.stack([fpx, fpy, fpz]), stress=st.reshape(3, 3, order="F"), dwrk=dw.value,
               nnlpp1=nn1[:nc1.value].copy(), ncount1=nc1.value)
    return out


def ref_ferrf_alchemy(s, charge, charge0, pairs, iz, fudge=1.0):
    R = ref()
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, dw = C.c_double(), C.c_double(), C.c_double()
    lst = _i(pairs)
    dummy = np.zeros(8)
    keep = [_i(s.ss_index), _f(charge), _f(charge0), _f(x), _f(y), _f(z)]
    R.ferrf_alchemy_(_p(keep[0]), _rd(s.alpha), _p(keep[1]), _p(keep[2]), _rd(fudge), _p(keep[3]), _p(keep[4]),
                     _p(keep[5]), _ri(iz), _p(lst), _ri(lst.size // 2), C.byref(a), C.byref(b), _p(fpx), _p(fpy),
                     _p(fpz), _ri(0), _p(dummy), _rd(0.0), _rd(0.0), _ri(0), C.byref(dw), _ri(n))
    return dict(slt=a.value, slv=b.value, dwrk=dw.value, f=np.stack([fpx, fpy, fpz]))


def ref_cself_alchemy(s, charge, charge0, kcut=1e9):
    R = ref()
    a, dw = C.c_double(), C.c_double()
    R.cself_alchemy_(_p(_i(s.ss_index)), _ri(s.natoms), _rd(s.alpha), _rd(kcut), _p(_f(charge)), _p(_f(charge0)),
                     C.byref(a), C.byref(dw))
    return a.value, dw.value


def ref_fft_pme_alchemy(s, charge, charge0, pressure=True, rkcut=1e9):
    """The reference fft_pme_alchemy (fftpack branch of do_pmesh_alchemy)."""
    R = ref()
    n = s.natoms
    R.refglue_pme_init(C.c_int(n), C.c_int(s.nfft[0]), C.c_int(s.nfft[1]), C.c_int(s.nfft[2]), C.c_int(s.order))
    cm = s.molecule_coms()
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    st = np.zeros(9)
    eer, dw = C.c_double(), C.c_double()
    arrs = [_f(s.xp0), _f(s.yp0), _f(s.zp0), _f(cm[0]), _f(cm[1]), _f(cm[2]), _f(charge), _f(charge0), _f(s.co),
            _f(s.oc)]
    atomp, grppt = _i(s.atomp), _i(s.grppt)

    def call():
        R.fft_pme_alchemy_(_ri(n), *[_p(a) for a in arrs[:8]], _p(arrs[8]), _p(arrs[9]), _rd(s.volume),
                           _rd(s.alpha), _ri(s.order), _ri(s.nfft[0]), _ri(s.nfft[1]), _ri(s.nfft[2]),
                           C.byref(eer), _p(fpx), _p(fpy), _p(fpz), _p(st), _p(atomp), _p(grppt),
                           _ri(1 if pressure else 0), _rd(rkcut), C.byref(dw))

    _big_stack(call)
    return dict(eer=eer.value, dwrk=dw.value, eer0=eer.value - dw.value, f=np.stack([fpx, fpy, fpz]),
                stress=st.reshape(3, 3, order="F"))


# ---------------------------------------------------------------- 1-4 pairs

FNB_E = ("uconf_slt", "uconf_slv", "ucoul_slt", "ucoul_slv")


def fnb14(s, int14, type14, fudge, lj_fudge, ewald=True, cut=True):
    """fnb14.f on the Cartesian coordinates of System s: pairs int14 (2, n)
    1-based, packed LJ indices type14 (1-based)."""
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    lst, t14 = _i(int14), _i(type14)
    lib().orc_fnb14(_p(_i(s.ss_index)), _p(_f(x)), _p(_f(y)), _p(_f(z)), _p(_f(s.charge)), C.c_int(n),
                    _p(_f(s.ecc12)), _p(_f(s.ecc6)), C.c_int(int(cut)), C.c_int(int(ewald)), C.c_double(s.alpha),
                    _p(lst), C.c_int(lst.size // 2), _p(t14), C.c_double(fudge), C.c_double(lj_fudge), _p(fpx),
                    _p(fpy), _p(fpz), *[C.byref(v) for v in e])
    return dict(zip(FNB_E, [v.value for v in e]), f=np.stack([fpx, fpy, fpz]))


def ref_fnb14(s, int14, type14, fudge, lj_fudge, ewald=True, cut=True):
    R = ref()
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    keep = [_i(s.ss_index), _f(x), _f(y), _f(z), _f(s.charge), _f(s.ecc12), _f(s.ecc6), _i(int14), _i(type14)]
    R.fnb14_(_p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]), _ri(n), _p(keep[5]), _p(keep[6]),
             _ri(int(cut)), _ri(int(ewald)), _rd(s.alpha), _p(keep[7]), _ri(keep[7].size // 2), _p(keep[8]),
             _rd(fudge), _rd(lj_fudge), _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e], _ri(1))
    return dict(zip(FNB_E, [v.value for v in e]), f=np.stack([fpx, fpy, fpz]))


def fptors(s, ltor, pota, potb, coords=None):
    """fptors.f on the Cartesian coordinates of System s (or coords = (x, y, z)):
    torsions ltor (4, n) 1-based, pota = signed multiplicity, potb = barrier.
    Returns utors_slt, utors_slv, f and status (1: a singular torsion was
    skipped, 2: unsupported multiplicity)."""
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    lt = _i(ltor)
    rc = lib().orc_fptors(_p(_i(s.ss_index)), _p(lt), C.c_int(lt.size // 4), _p(_f(x)), _p(_f(y)), _p(_f(z)),
                          _p(_f(pota)), _p(_f(potb)), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]), status=rc)


def ref_fptors(s, ltor, pota, potb, coords=None):
    R = ref()
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    keep = [_i(s.ss_index), _i(ltor), _f(x), _f(y), _f(z), _f(pota), _f(potb)]
    R.fptors_(_p(keep[0]), _p(keep[1]), _ri(keep[1].size // 4), _p(keep[2]), _p(keep[3]), _p(keep[4]), _p(keep[5]),
              _p(keep[6]), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _ri(n), _ri(1))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]))



def fitors(s, litr, pota, potb, potc, coords=None):
    """fitors.f on the Cartesian coordinates of System s (or coords): improper
    torsions litr (4, n) 1-based; potc > 0 selects the harmonic form (pota the
    rest dihedral, potb the constant), else pota is the signed multiplicity.
    Returns utors_slt, utors_slv, f and status (2: unsupported multiplicity)."""
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    lt = _i(litr)
    rc = lib().orc_fitors(_p(_i(s.ss_index)), _p(lt), C.c_int(lt.size // 4), _p(_f(x)), _p(_f(y)), _p(_f(z)),
                          _p(_f(pota)), _p(_f(potb)), _p(_f(potc)), C.byref(a), C.byref(b), _p(fpx), _p(fpy),
                          _p(fpz))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]), status=rc)


def ref_fitors(s, litr, pota, potb, potc, coords=None):
    R = ref()
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    keep = [_i(s.ss_index), _i(litr), _f(x), _f(y), _f(z), _f(pota), _f(potb), _f(potc)]
    R.fitors_(_p(keep[0]), _p(keep[1]), _ri(keep[1].size // 4), _p(keep[2]), _p(keep[3]), _p(keep[4]), _p(keep[5]),
              _p(keep[6]), _p(keep[7]), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _ri(n), _ri(1))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]))

# ---------------------------------------------------------------- REM twins of the bonded terms

def _flags(v, n):
    f = np.zeros(n, np.int32)
    f[: len(v)] = np.asarray(v, dtype=bool)
    return f


def fnb14_rem(s, int14, type14, fudge, lj_fudge, remint14, f2, ewald=True, cut=True):
    """fnb14_rem.f: fnb14 with per-pair flags remint14 (energy to seg_ene, force x f2)."""
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    seg = C.c_double()
    lst, t14 = _i(int14), _i(type14)
    fl = _flags(remint14, lst.size // 2)
    lib().orc_fnb14_rem(_p(_i(s.ss_index)), _p(_f(x)), _p(_f(y)), _p(_f(z)), _p(_f(s.charge)), C.c_int(n),
                        _p(_f(s.ecc12)), _p(_f(s.ecc6)), C.c_int(int(cut)), C.c_int(int(ewald)), C.c_double(s.alpha),
                        _p(lst), C.c_int(lst.size // 2), _p(t14), C.c_double(fudge), C.c_double(lj_fudge), _p(fpx),
                        _p(fpy), _p(fpz), *[C.byref(v) for v in e], _p(fl), C.c_double(f2), C.byref(seg))
    return dict(zip(FNB_E, [v.value for v in e]), f=np.stack([fpx, fpy, fpz]), seg_ene=seg.value)


def ref_fnb14_rem(s, int14, type14, fudge, lj_fudge, remint14, f2, ewald=True, cut=True):
    R = ref()
    R.refglue_set_rem_factor(C.c_int(2), C.c_double(float(f2)))
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    seg = C.c_double()
    keep = [_i(s.ss_index), _f(x), _f(y), _f(z), _f(s.charge), _f(s.ecc12), _f(s.ecc6), _i(int14), _i(type14)]
    fl = _flags(remint14, keep[7].size // 2)
    R.fnb14_rem_(_p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]), _ri(n), _p(keep[5]), _p(keep[6]),
                 _ri(int(cut)), _ri(int(ewald)), _rd(s.alpha), _p(keep[7]), _ri(keep[7].size // 2), _p(keep[8]),
                 _rd(fudge), _rd(lj_fudge), _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e], _p(fl), _ri(1),
                 C.byref(seg))
    return dict(zip(FNB_E, [v.value for v in e]), f=np.stack([fpx, fpy, fpz]), seg_ene=seg.value)


def fptors_rem(s, ltor, pota, potb, remltor, f2, coords=None):
    """fptors_rem.f: fptors with per-torsion flags remltor."""
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, seg = C.c_double(), C.c_double(), C.c_double()
    lt = _i(ltor)
    fl = _flags(remltor, lt.size // 4)
    rc = lib().orc_fptors_rem(_p(_i(s.ss_index)), _p(lt), C.c_int(lt.size // 4), _p(_f(x)), _p(_f(y)), _p(_f(z)),
                              _p(_f(pota)), _p(_f(potb)), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz),
                              _p(fl), C.c_double(f2), C.byref(seg))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]), status=rc, seg_ene=seg.value)


def ref_fptors_rem(s, ltor, pota, potb, remltor, f2, coords=None):
    R = ref()
    R.refglue_set_rem_factor(C.c_int(2), C.c_double(float(f2)))
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, seg = C.c_double(), C.c_double(), C.c_double()
    keep = [_i(s.ss_index), _i(ltor), _f(x), _f(y), _f(z), _f(pota), _f(potb)]
    fl = _flags(remltor, keep[1].size // 4)
    R.fptors_rem_(_p(keep[0]), _p(keep[1]), _ri(keep[1].size // 4), _p(keep[2]), _p(keep[3]), _p(keep[4]),
                  _p(keep[5]), _p(keep[6]), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _p(fl), _ri(1),
                  C.byref(seg))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]), seg_ene=seg.value)


def fitors_rem(s, litr, pota, potb, potc, remitor, f2, coords=None):
    """fitors_rem.f: fitors with per-term flags remitor."""
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, seg = C.c_double(), C.c_double(), C.c_double()
    lt = _i(litr)
    fl = _flags(remitor, lt.size // 4)
    rc = lib().orc_fitors_rem(_p(_i(s.ss_index)), _p(lt), C.c_int(lt.size // 4), _p(_f(x)), _p(_f(y)), _p(_f(z)),
                              _p(_f(pota)), _p(_f(potb)), _p(_f(potc)), C.byref(a), C.byref(b), _p(fpx), _p(fpy),
                              _p(fpz), _p(fl), C.c_double(f2), C.byref(seg))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]), status=rc, seg_ene=seg.value)


def ref_fitors_rem(s, litr, pota, potb, potc, remitor, f2, coords=None):
    R = ref()
    R.refglue_set_rem_factor(C.c_int(2), C.c_double(float(f2)))
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, seg = C.c_double(), C.c_double(), C.c_double()
    keep = [_i(s.ss_index), _i(litr), _f(x), _f(y), _f(z), _f(pota), _f(potb), _f(potc)]
    fl = _flags(remitor, keep[1].size // 4)
    R.fitors_rem_(_p(keep[0]), _p(keep[1]), _ri(keep[1].size // 4), _p(keep[2]), _p(keep[3]), _p(keep[4]),
                  _p(keep[5]), _p(keep[6]), _p(keep[7]), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _p(fl),
                  _ri(n), _ri(1), C.byref(seg))
    return dict(utors_slt=a.value, utors_slv=b.value, f=np.stack([fpx, fpy, fpz]), seg_ene=seg.value)


def fpbond(s, lbnd, pota, potb, coords=None):
    """fpbond.f on the Cartesian coordinates of System s (or coords): bonds
    lbnd (2, n) 1-based, pota = rest length, potb = force constant."""
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    lb = _i(lbnd)
    lib().orc_fpbond(_p(_i(s.ss_index)), _p(lb), C.c_int(lb.size // 2), _p(_f(x)), _p(_f(y)), _p(_f(z)),
                     _p(_f(pota)), _p(_f(potb)), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz))
    return dict(ubond_slt=a.value, ubond_slv=b.value, f=np.stack([fpx, fpy, fpz]))


def ref_fpbond(s, lbnd, pota, potb, coords=None):
    R = ref()
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    keep = [_i(s.ss_index), _i(lbnd), _f(x), _f(y), _f(z), _f(pota), _f(potb)]
    R.fpbond_(_p(keep[0]), _p(keep[1]), _ri(keep[1].size // 2), _p(keep[2]), _p(keep[3]), _p(keep[4]), _p(keep[5]),
              _p(keep[6]), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _ri(0), _ri(n), _ri(1))
    return dict(ubond_slt=a.value, ubond_slv=b.value, f=np.stack([fpx, fpy, fpz]))


def fpbend(s, lbndg, pots, coords=None):
    """fpbend.f on the Cartesian coordinates of System s (or coords): bends
    lbndg (3, n) 1-based, pots = (pota, potb, potc, potd, pote, potf, potg) as
    fpbend receives them (rest angle, constant, Urey-Bradley rest distance and
    constant, coupling parameters)."""
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    lb = _i(lbndg)
    pk = [_f(v) for v in pots]
    lib().orc_fpbend(_p(_i(s.ss_index)), _p(lb), C.c_int(lb.size // 3), _p(_f(x)), _p(_f(y)), _p(_f(z)),
                     *[_p(v) for v in pk], C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz))
    return dict(ubend_slt=a.value, ubend_slv=b.value, f=np.stack([fpx, fpy, fpz]))


def ref_fpbend(s, lbndg, pots, coords=None):
    R = ref()
    x, y, z = coords if coords is not None else s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b = C.c_double(), C.c_double()
    keep = [_i(s.ss_index), _i(lbndg), _f(x), _f(y), _f(z)] + [_f(v) for v in pots]
    R.fpbend_(_p(keep[0]), _p(keep[1]), _ri(keep[1].size // 3), _p(keep[2]), _p(keep[3]), _p(keep[4]),
              *[_p(v) for v in keep[5:]], C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _ri(n), _ri(1))
    return dict(ubend_slt=a.value, ubend_slv=b.value, f=np.stack([fpx, fpy, fpz]))


def ref_int_corr_erf_spline(rmin, rmax, nbin, alpha, kcut):
    """The reference's ERF_CORR table builder: (erf_arr_corr (4, nbin + 1), tau)."""
    tab = np.zeros((4, nbin + 1), order="F")
    tau = np.zeros(nbin + 1)
    ref().int_corr_erf_spline_(_rd(rmin), _rd(rmax), _ri(nbin), _rd(alpha), _rd(kcut), _p(tab), _p(tau))
    return tab, tau


def ref_erf_corr_cutoff(oc, nfft):
    rk, dl = C.c_double(), C.c_double()
    ref().erf_corr_cutoff_(_p(_f(oc)), C.byref(dl), C.byref(rk), _ri(nfft[0]), _ri(nfft[1]), _ri(nfft[2]))
    return rk.value


# ---------------------------------------------------------------- alchemical intramolecular terms

def ref_decide_force(rinn, alpha):
    """The reference's decide_force (mts_forces.f:212-241)."""
    out = C.c_int32(0)
    ref().decide_force_(_rd(rinn), _rd(alpha), C.byref(out))
    return bool(out.value)


def ferrf_alchemy_add(s, lambdaq, lambdaq0, listqq, intramol=False, erf=None, charge=None):
    """ferrf_alchemy_add.f on the Cartesian coordinates of System s: listqq
    (3, n) = (ia, ib, facQ), 1-based; per-atom lambdaq, lambdaq0; the created
    charges (s.charge unless given); erf = (table, delew, rlew) selects the
    ERF_CORR branch."""
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, dw = C.c_double(), C.c_double(), C.c_double()
    lst = _i(listqq)
    tab = _f(erf[0]) if erf is not None else np.zeros(4)
    q = s.charge if charge is None else charge
    lib().orc_ferrf_alchemy_add(_p(_i(s.ss_index)), C.c_double(s.alpha), _p(_f(q)), _p(_f(lambdaq)),
                                _p(_f(lambdaq0)), _p(_f(x)), _p(_f(y)), _p(_f(z)), _p(lst), C.c_int(lst.size // 3),
                                C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), C.c_int32(0 if erf is None else 1),
                                _p(tab), C.c_double(erf[1] if erf is not None else 0.0),
                                C.c_double(erf[2] if erf is not None else 0.0), C.byref(dw), C.c_int(int(intramol)))
    return dict(slt=a.value, slv=b.value, dwrk=dw.value, f=np.stack([fpx, fpy, fpz]))


def ref_ferrf_alchemy_add(s, lambdaq, lambdaq0, listqq, intramol=False, erf=None, charge=None):
    R = ref()
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    a, b, dw = C.c_double(), C.c_double(), C.c_double()
    lst = _i(listqq)
    tab = _f(erf[0]) if erf is not None else np.zeros(8)
    q = s.charge if charge is None else charge
    keep = [_i(s.ss_index), _f(q), _f(lambdaq), _f(lambdaq0), _f(x), _f(y), _f(z)]
    R.ferrf_alchemy_add_(_p(keep[0]), _rd(s.alpha), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]), _p(keep[5]),
                         _p(keep[6]), _p(lst), _ri(lst.size // 3), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz),
                         _ri(0 if erf is None else 1), _p(tab), _rd(erf[1] if erf is not None else 0.0),
                         _rd(erf[2] if erf is not None else 0.0), C.byref(dw), _ri(n), _ri(int(intramol)))
    return dict(slt=a.value, slv=b.value, dwrk=dw.value, f=np.stack([fpx, fpy, fpz]))


def fnb14_alchemy(s, int14, type14, fudge, lj_fudge, lams, ewald=True, cut=True):
    """fnb14_alchemy.f on the Cartesian coordinates of System s; lams =
    (lambda, lambdaq, lambda0, lambdaq0) per atom."""
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    dw = C.c_double()
    lst, t14 = _i(int14), _i(type14)
    lf = [_f(v) for v in lams]
    lib().orc_fnb14_alchemy(_p(_i(s.ss_index)), _p(_f(x)), _p(_f(y)), _p(_f(z)), _p(_f(s.charge)), C.c_int(n),
                            _p(_f(s.ecc12)), _p(_f(s.ecc6)), C.c_int(int(cut)), C.c_int(int(ewald)),
                            C.c_double(s.alpha), _p(lst), C.c_int(lst.size // 2), _p(t14), C.c_double(fudge),
                            C.c_double(lj_fudge), _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e],
                            *[_p(v) for v in lf], C.byref(dw))
    return dict(zip(FNB_E, [v.value for v in e]), dwrk=dw.value, f=np.stack([fpx, fpy, fpz]))


def ref_fnb14_alchemy(s, int14, type14, fudge, lj_fudge, lams, ewald=True, cut=True):
    R = ref()
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    dw = C.c_double(0.0)
    keep = [_i(s.ss_index), _f(x), _f(y), _f(z), _f(s.charge), _f(s.ecc12), _f(s.ecc6), _i(int14), _i(type14)]
    lf = [_f(v) for v in lams]
    R.fnb14_alchemy_(_p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]), _ri(n), _p(keep[5]),
                     _p(keep[6]), _ri(int(cut)), _ri(int(ewald)), _rd(s.alpha), _p(keep[7]), _ri(keep[7].size // 2),
                     _p(keep[8]), _rd(fudge), _rd(lj_fudge), _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e],
                     *[_p(v) for v in lf], C.byref(dw), _ri(1))
    return dict(zip(FNB_E, [v.value for v in e]), dwrk=dw.value, f=np.stack([fpx, fpy, fpz]))


def fnbgrp_alchemy(s, ingrp, typ, ewald, lambdaq):
    """fnbgrp_alchemy.f: the arithmetic of fnbgrp.f (the alphal0 it derives
    from lambdaq at :152-157 is never read), so the oracle is fnbgrp()."""
    del lambdaq
    return fnbgrp(s, ingrp, typ, ewald)


def ref_fnbgrp_alchemy(s, ingrp, typ, ewald, lambdaq):
    R = ref()
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
    e = [C.c_double() for _ in range(4)]
    keep = [_i(s.ss_index), _f(x), _f(y), _f(z), _f(s.charge), _i(typ), _f(s.ecc12), _f(s.ecc6), _i(ingrp),
            _f(lambdaq)]
    R.fnbgrp_alchemy_(*[_p(k) for k in keep[:8]], _ri(int(ewald)), _rd(s.alpha), _p(keep[8]),
                      _ri(keep[8].size // 2), _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e], _p(keep[9]),
                      _ri(1))
    return dict(uconf_slt=e[0].value, uconf_slv=e[1].value, ucoul_slt=e[2].value, ucoul_slv=e[3].value,
                f=np.stack([fpx, fpy, fpz]))


# ---------------------------------------------------------------- constraints and the integrator step

def constraint_tables(first, lens, mass):
    """ORAC's constraint arrays for rigid 3-site units: first (1-based first
    atoms), lens (n, 3) lengths of (0,1), (0,2), (1,2).  Returns cnst (2, 3n)
    1-based, dss (squared lengths), coeff (bcnstp.f:79-80) and cnst_protl
    (per molecule [3, k, k+1, k+2])."""
    first = np.asarray(first, np.int64)
    n = len(first)
    pairs = ((0, 1), (0, 2), (1, 2))
    cnst = np.zeros((2, 3 * n), np.int32)
    for c, (i, j) in enumerate(pairs):
        cnst[0, c::3] = first + i
        cnst[1, c::3] = first + j
    dss = (np.asarray(lens, np.float64).reshape(n, 3) ** 2).ravel()
    m = np.asarray(mass, np.float64)
    ma, mb = m[cnst[0] - 1], m[cnst[1] - 1]
    coeff = 2.0 * (ma + mb) / (ma * mb)
    prot = np.zeros((n, 4), np.int32)
    prot[:, 0] = 3
    prot[:, 1:] = 1 + np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return cnst, dss, coeff, prot.ravel()


def _cons_args(x, v, cnst, dss, coeff, mass, prot):
    keep = [np.array(a, dtype=np.float64, order="C") for a in (x[0], x[1], x[2], v[0], v[1], v[2])]
    tabs = [_i(cnst), _f(dss), _f(coeff), _f(mass), _i(prot)]
    return keep, tabs


def rattle_verlet(dt, x1, x0, v, cnst, dss, coeff, mass, prot, mim_lim=20):
    """rattle_verlet.f: positions x1 (3, n) moved onto the constraints along
    the bonds of x0, velocities v corrected by the displacement / dt.
    Returns (x1, v, dnit increment, status)."""
    keep, (c, d, cf, m, p) = _cons_args(x1, v, cnst, dss, coeff, mass, prot)
    x0 = [_f(a) for a in x0]
    dnit = C.c_double(0.0)
    n = len(keep[0])
    rc = lib().orc_rattle_verlet(C.c_double(dt), *[_p(a) for a in keep[:3]], *[_p(a) for a in x0],
                                 *[_p(a) for a in keep[3:]], C.c_int(n), _p(c), _p(d), _p(cf), C.c_int(len(d)),
                                 _p(m), C.c_int(_nprot(p)),
                                 _p(p), C.c_int(mim_lim), C.byref(dnit))
    return np.stack(keep[:3]), np.stack(keep[3:]), dnit.value, rc


def _nprot(prot):
    """Molecules in a cnst_protl array."""
    n, k = 0, 0
    while k < len(prot):
        k += int(prot[k]) + 1
        n += 1
    return n


def rattle_correc(x, v, cnst, dss, coeff, mass, prot, mim_lim=20):
    """rattle_correc.f: velocities v made orthogonal to the bonds at x.
    Returns (v, dnit increment, status)."""
    keep, (c, d, cf, m, p) = _cons_args(x, v, cnst, dss, coeff, mass, prot)
    dnit = C.c_double(0.0)
    n = len(keep[0])
    rc = lib().orc_rattle_correc(C.c_double(0.0), *[_p(a) for a in keep[:3]], *[_p(a) for a in keep[3:]], C.c_int(n),
                                 _p(c), _p(d), _p(cf), C.c_int(len(d)), _p(m), C.c_int(_nprot(p)), _p(p),
                                 C.c_int(mim_lim), C.byref(dnit))
    return np.stack(keep[3:]), dnit.value, rc


def ref_rattle_verlet(dt, x1, x0, v, cnst, dss, coeff, mass, prot, mim_lim=20):
    """The reference's rattle_verlet (compiled in place), same arguments."""
    R = ref()
    keep, (c, d, cf, m, p) = _cons_args(x1, v, cnst, dss, coeff, mass, prot)
    x0 = [_f(a) for a in x0]
    dnit, iret = C.c_double(0.0), C.c_int(0)
    err = C.create_string_buffer(80)
    n = len(keep[0])
    R.rattle_verlet_(C.byref(C.c_double(dt)), *[_p(a) for a in keep[:3]], *[_p(a) for a in x0],
                     *[_p(a) for a in keep[3:]], _ri(n), _p(c), _p(d), _p(cf), _ri(len(d)), _p(m), C.byref(dnit),
                     _ri(_nprot(p)), _p(p), _ri(mim_lim), C.byref(iret), err, C.c_size_t(80))
    return np.stack(keep[:3]), np.stack(keep[3:]), dnit.value, iret.value


def ref_rattle_correc(x, v, cnst, dss, coeff, mass, prot, mim_lim=20):
    R = ref()
    keep, (c, d, cf, m, p) = _cons_args(x, v, cnst, dss, coeff, mass, prot)
    dnit, iret = C.c_double(0.0), C.c_int(0)
    err = C.create_string_buffer(80)
    n = len(keep[0])
    R.rattle_correc_(C.byref(C.c_double(0.0)), *[_p(a) for a in keep[:3]], *[_p(a) for a in keep[3:]], _ri(n),
                     _p(c), _p(d), _p(cf), _ri(len(d)), _p(m), C.byref(dnit), _ri(_nprot(p)), _p(p), _ri(mim_lim),
                     C.byref(iret), err, C.c_size_t(80))
    return np.stack(keep[3:]), dnit.value, iret.value


def correc(v, f, mass, dt):
    """correc.f:1-46: v += f*0.5*dt/mass (a half kick), in the reference's
    association order."""
    return v + f * 0.5 * dt / mass


def verlet(x, v, f, mass, dt):
    """verlet.f:1-52: x += dt*v + f*0.5*dt^2/mass, v += f*0.5*dt/mass."""
    ts2 = dt * dt
    return x + dt * v + f * 0.5 * ts2 / mass, v + f * 0.5 * dt / mass


def ref_correc(v, f, mass, dt):
    R = ref()
    n = v.shape[1]
    vv = [np.array(a, dtype=np.float64) for a in v]
    ff = [_f(a) for a in f]
    R.correc_(*[_p(a) for a in vv], *[_p(a) for a in ff], _p(_f(mass)), _ri(n), C.byref(C.c_double(dt)))
    return np.stack(vv)


def ref_verlet(x, v, f, mass, dt):
    R = ref()
    n = x.shape[1]
    xx = [np.array(a, dtype=np.float64) for a in x]
    vv = [np.array(a, dtype=np.float64) for a in v]
    ff = [_f(a) for a in f]
    R.verlet_(_p(_f(mass)), _ri(n), *[_p(a) for a in xx], *[_p(a) for a in vv], *[_p(a) for a in ff],
              C.byref(C.c_double(dt)))
    return np.stack(xx), np.stack(vv)


def md_steps(x, v, mass, nsteps, dt, ninner, slow, fast, cons):
    """The device integrator's step, restated: two levels of ORAC's r-RESPA
    (mtsmd.f), the slow m level (full nonbonded forces, once per dt) around
    ninner fast n0 substeps of h = dt/ninner, in mtsmd.f's order:

        correc(Fs, dt); rattle_correc                       (:2598-2604)
        ninner times: verlet(Ff, h); rattle_verlet(h)       (:2663-2683)
                      Ff = fast(x); correc(Ff, h); rattle_correc  (:2843-2884)
        Fs = slow(x); correc(Fs, dt); rattle_correc         (:3030-3146)

    x, v (3, n); slow(x) and fast(x) return (3, n) forces at Cartesian
    positions x; cons = constraint_tables(...).  Returns (x, v, failures).

    Its pieces (correc, verlet, rattle_verlet, rattle_correc) are pinned bit
    for bit against the reference; the order of the calls is restated from
    mtsmd.f, which is the caller and is not compiled here: parity unpinned
    for that order (it is checked by time reversibility, NVE drift and the
    end-of-step invariants: RATTLE last, the last drift SHAKEn)."""
    cnst, dss, coeff, prot = cons
    h = dt / ninner
    fs, ff = slow(x), fast(x)
    fails = 0
    for _ in range(nsteps):
        v = correc(v, fs, mass, dt)
        v, _, rc = rattle_correc(x, v, cnst, dss, coeff, mass, prot)
        fails += rc != 0
        for _ in range(ninner):
            xo = x
            x, v = verlet(x, v, ff, mass, h)
            x, v, _, rc = rattle_verlet(h, x, xo, v, cnst, dss, coeff, mass, prot)
            fails += rc != 0
            ff = fast(x)
            v = correc(v, ff, mass, h)
            v, _, rc = rattle_correc(x, v, cnst, dss, coeff, mass, prot)
            fails += rc != 0
        fs = slow(x)
        v = correc(v, fs, mass, dt)
        v, _, rc = rattle_correc(x, v, cnst, dss, coeff, mass, prot)
        fails += rc != 0
    return x, v, fails


def harmonic_forces(x, pairs, k, r0):
    """Forces of E = k (r - r0)^2 over 1-based pairs (2, m) at Cartesian x
    (3, n): the device integrator's fast terms."""
    f = np.zeros_like(x)
    if pairs.shape[1] == 0:
        return f
    i, j = pairs[0] - 1, pairs[1] - 1
    d = x[:, i] - x[:, j]
    r = np.sqrt((d * d).sum(axis=0))
    c = -2.0 * k * (r - r0) / r
    for a in range(3):
        np.add.at(f[a], i, c * d[a])
        np.add.at(f[a], j, -c * d[a])
    return f


# ---------------------------------------------------------------- the reference's OpenMP build (CPU baseline)

_ref_omp = None
_omp_state = {}


def ref_omp_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref_omp.so"))


def ref_omp(nthr):
    """The reference's hot path as ORAC's OpenMP build compiles it
    (_ref/libref_omp.so), set up for nthr threads (at most 16, the
    reference's mcor3).  Returns (library, mpp8, m8t)."""
    global _ref_omp
    if not 1 <= nthr <= 16:
        raise ValueError("the reference's OpenMP build runs 1 to 16 threads (omp_integr mcor3)")
    if _ref_omp is None:
        # the per-thread private arrays of force.f need more than the default
        # worker stack; read when the OpenMP runtime starts
        os.environ.setdefault("OMP_STACKSIZE", "512M")
        _ref_omp = C.CDLL(os.path.join(HERE, "_ref", "libref_omp.so"))
    if _omp_state.get("nthr") != nthr:
        mpp8, m8t = C.c_int(), C.c_int()
        _ref_omp.refglue_omp_setup(C.c_int(nthr), C.byref(mpp8), C.byref(m8t))
        _omp_state.update(nthr=nthr, mpp8=mpp8.value, m8t=m8t.value)
    return _ref_omp, _omp_state["mpp8"], _omp_state["m8t"]


def ref_omp_evaluation(s, corr_pairs, nthr, fa=None, nsample=1, budget_s=None):
    """One full evaluation (or nsample of them, stopping early after budget_s
    seconds) through the reference's OpenMP build on nthr threads, in ORAC's
    layouts: lc_index, lc_list into per-thread lists nnlpp(mpp8, nthr), force
    with the per-thread contact maps (split_map), fft_pme (fftpack branch)
    and ferrf over corr_pairs (fudge 1).  Returns (forces (3, N), evaluations
    done, seconds)."""
    import time as _t
    R, mpp8, m8t = ref_omp(nthr)
    fa = fa or ForceArgs(s)
    n = s.natoms
    rc = s.rcut + s.rtol + s.rnei
    nc = int(s.L // 3.0)
    # contact maps per thread, once (ORAC splits them at start-up)
    mapnlt = np.zeros(m8t * nthr, dtype=np.int32)
    R.refglue_omp_split_map(C.c_int(n), _p(fa.mapnl), _p(fa.grppt), C.c_int(s.ngrp), _p(mapnlt))
    keep = [_i(s.ss_index), _f(s.xp0), _f(s.yp0), _f(s.zp0), _f(s.charge), _i(s.nbtype), _i(s.atomg),
            _i(s.atomp), _f(s.co), _f(s.ecc12), _f(s.ecc6), _f(s.pmass), _f(s.mass),
            np.zeros(s.ngrp, dtype=np.int32)]
    (ss, x, y, z, q, nbt, atg, atp, co, e12, e6, pm, ms, groupp) = keep
    nnlpp = np.zeros(mpp8 * nthr, dtype=np.int32)
    nnlpp1 = np.zeros(mpp8 * nthr, dtype=np.int32)
    ncount1 = np.zeros(nthr, dtype=np.int32)
    nind = np.zeros(2, dtype=np.int32)
    indmax = 2500
    ii, jj, kk = [np.zeros(2 * indmax, dtype=np.int32) for _ in range(3)]
    cm = s.molecule_coms()
    parr = [_f(s.xp0), _f(s.yp0), _f(s.zp0), _f(cm[0]), _f(cm[1]), _f(cm[2]), _f(s.charge), _f(s.co), _f(s.oc)]
    cx, cy, cz = [_f(a) for a in s.cartesian()]
    lst = _i(corr_pairs)
    tab = np.zeros(8)
    f = {}

    def one():
        R.lc_index_(_ri(indmax), _ri(nc), _ri(nc), _ri(nc), _p(nind), _p(ii), _p(jj), _p(kk), _rd(rc), _p(co))
        R.lc_list_(_ri(nc), _ri(nc), _ri(nc), _p(nind), _p(ii), _p(jj), _p(kk), _rd(rc), _p(co), _p(fa.xpg),
                   _p(fa.ypg), _p(fa.zpg), _ri(s.ngrp), _p(nnlpp), _ri(6))
        fpx, fpy, fpz = [np.zeros(n) for _ in range(3)]
        st = np.zeros(9)
        e = [C.c_double() for _ in range(6)]
        rti = C.c_double(0.0)
        R.force_(_p(ss), _p(x), _p(y), _p(z), _p(fa.xpg), _p(fa.ypg), _p(fa.zpg), _p(q), _p(nbt), _p(fa.tt),
                 _ri(s.ntypes), _ri(n), _p(atg), _p(fa.xpcm), _p(fa.ypcm), _p(fa.zpcm), _p(groupp), _p(atp),
                 _p(co), _p(e12), _p(e6), _rd(s.alpha), _p(mapnlt), _ri(s.ngrp), _p(fa.grppt),
                 *[C.byref(v) for v in e], _p(fpx), _p(fpy), _p(fpz), _p(st), _p(nnlpp), _p(nnlpp1), _rd(0.0),
                 _rd(0.0), _rd(s.rcut), C.byref(rti), _rd(s.rtol), _p(pm), _ri(1), _p(ncount1), _p(ms),
                 _rd(1.0e20), _ri(0), _ri(1))
        gx, gy, gz = [np.zeros(n) for _ in range(3)]
        st2 = np.zeros(9)
        eer = C.c_double()
        R.refglue_omp_pme_init(C.c_int(n), C.c_int(s.nfft[0]), C.c_int(s.nfft[1]), C.c_int(s.nfft[2]),
                               C.c_int(s.order))
        R.fft_pme_(_ri(n), *[_p(a) for a in parr[:7]], _p(parr[7]), _p(parr[8]), _rd(s.volume), _rd(s.alpha),
                   _ri(s.order), _ri(s.nfft[0]), _ri(s.nfft[1]), _ri(s.nfft[2]), C.byref(eer), _p(gx), _p(gy),
                   _p(gz), _p(st2), _p(atp), _p(fa.grppt), _ri(1), _rd(1e9))
        hx, hy, hz = [np.zeros(n) for _ in range(3)]
        a, b = C.c_double(), C.c_double()
        if lst.size:
            R.ferrf_(_p(ss), _rd(s.alpha), _p(q), _rd(1.0), _p(cx), _p(cy), _p(cz), _ri(1), _p(lst),
                     _ri(lst.size // 2), C.byref(a), C.byref(b), _p(hx), _p(hy), _p(hz), _ri(0), _p(tab), _rd(0.0),
                     _rd(0.0), _ri(0), _ri(n))
        f["f"] = np.stack([fpx + gx + hx, fpy + gy + hy, fpz + gz + hz])

    def run():
        R.refglue_omp_threads(C.c_int(nthr))  # the team size is a setting of the calling thread
        t0 = _t.perf_counter()
        done = 0
        while done < nsample:
            one()
            done += 1
            if budget_s is not None and _t.perf_counter() - t0 > budget_s:
                break
        return done, _t.perf_counter() - t0

    done, secs = _big_stack(run)
    return f["f"], done, secs


def md_steps_shells(x, v, mass, nsteps, dt, nmid, ninner, slow, mid, fast, cons):
    """md_steps with two nonbonded shells: three levels of mtsmd.f, the outer
    shell (slow: outer-shell direct forces, PME, corrections) over dt, the
    inner shell (mid) nmid times per dt, the fast terms ninner times per
    dt/nmid, in mtsmd.f's order (its m level inside h, :2598-3146):

        correc(Fs, dt); rattle_correc
        nmid times: correc(Fm, hm); rattle_correc
                    ninner times: verlet(Ff, h); rattle_verlet; Ff = fast(x); correc(Ff, h); rattle_correc
                    Fm = mid(x) (and Fs = slow(x) after the last); correc(Fm, hm); rattle_correc
        correc(Fs, dt); rattle_correc

    Returns (x, v, failures)."""
    cnst, dss, coeff, prot = cons
    hm = dt / nmid
    h = hm / ninner
    fs, fm, ff = slow(x), mid(x), fast(x)
    fails = 0

    def rc_(xx, vv):
        nonlocal fails
        vv, _, rc = rattle_correc(xx, vv, cnst, dss, coeff, mass, prot)
        fails += rc != 0
        return vv

    for _ in range(nsteps):
        v = rc_(x, correc(v, fs, mass, dt))
        for im in range(nmid):
            v = rc_(x, correc(v, fm, mass, hm))
            for _ in range(ninner):
                xo = x
                x, v = verlet(x, v, ff, mass, h)
                x, v, _, rc = rattle_verlet(h, x, xo, v, cnst, dss, coeff, mass, prot)
                fails += rc != 0
                ff = fast(x)
                v = rc_(x, correc(v, ff, mass, h))
            fm = mid(x)
            if im == nmid - 1:
                fs = slow(x)
            v = rc_(x, correc(v, fm, mass, hm))
        v = rc_(x, correc(v, fs, mass, dt))
    return x, v, fails
