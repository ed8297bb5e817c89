"""Python mirror of ORAC's nonbonded routines over the gfx950 C ABI.

``Context`` methods carry the reference routines' names and argument lists
(``fft_pme_init``, ``fft_pme``, ``lc_index``, ``lc_list``, ``force``,
``ferrf``, ``cself``, ``fnbgrp``, ``decide_force``).  Arrays are numpy arrays
in the reference's layouts (Fortran column-major, 1-based index arrays);
output arrays are updated in place, as the Fortran callees update their
arguments, and scalar outputs are returned.  A nonzero status from the
library raises ``OracHipError`` (the reference's fatal ``xerror`` path).

``ResidentSystem`` drives the device-resident path (topology, coordinates,
lists and grids kept in HBM); ``Batch`` evaluates several resident systems
(replicas, windows, requests) with one kernel launch per stage.

The library must be built (``orac_hip.build``); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

_lib = None


class OracHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"orac_hip status {code}: {msg}")
        self.code = code


def library_path():
    return _build.LIB


def lib():
    """Load liborac_hip.so; raises if it was not built."""
    global _lib
    if _lib is None:
        path = _build.LIB
        if not os.path.exists(path):
            raise OracHipError(5, f"{path} is missing: build it with orac_hip.build.build()")
        _lib = C.CDLL(path)
        for name in ("orac_hip_create", "orac_hip_last_error"):
            getattr(_lib, name).restype = C.c_int
    return _lib


EXPORTS = [
    "orac_hip_create", "orac_hip_destroy", "orac_hip_last_error", "orac_hip_invalidate", "orac_hip_device_name",
    "orac_hip_probe_copy",
    "orac_hip_fft_pme_init", "orac_hip_fft_pme", "orac_hip_fft_pme_noalloc", "orac_hip_lc_index", "orac_hip_lc_list", "orac_hip_force",
    "orac_hip_force_rem", "orac_hip_fptors_rem", "orac_hip_fitors_rem", "orac_hip_fnb14_rem",
    "orac_hip_force_alchemy", "orac_hip_force_alchemy_shift", "orac_hip_fptors", "orac_hip_fpbond", "orac_hip_fitors", "orac_hip_fpbend", "orac_hip_fnb14",
    "orac_hip_fft_pme_alchemy", "orac_hip_ferrf_alchemy", "orac_hip_cself_alchemy", "orac_hip_int_corr_erf_spline",
    "orac_hip_erf_corr_cutoff",
    "orac_hip_decide_force", "orac_hip_ferrf", "orac_hip_cself", "orac_hip_fnbgrp", "orac_hip_system_create",
    "orac_hip_system_destroy", "orac_hip_system_set_coords", "orac_hip_system_device_ptrs", "orac_hip_system_eval",
    "orac_hip_system_results", "orac_hip_system_profile", "orac_hip_system_stage_ms", "orac_hip_system_bytes",
    "orac_hip_sync", "orac_hip_stream", "orac_hip_batch_create", "orac_hip_batch_destroy", "orac_hip_batch_eval",
    "orac_hip_batch_profile", "orac_hip_batch_stage_ms", "orac_hip_batch_set_tuning", "orac_hip_batch_get_tuning",
    "orac_hip_batch_autotune", "orac_hip_batch_tune_log", "orac_hip_md_setup", "orac_hip_md_set_state", "orac_hip_md_get_state",
    "orac_hip_batch_md_run", "orac_hip_batch_md_run_shells", "orac_hip_system_set_shells",
    "orac_hip_batch_md_relax", "orac_hip_batch_md_scale_velocities",
    "orac_hip_system_md_energies", "orac_hip_system_set_alchemy", "orac_hip_system_set_lambda",
    "orac_hip_system_alchemy_results", "orac_hip_ferrf_alchemy_add", "orac_hip_fnb14_alchemy",
    "orac_hip_fnbgrp_alchemy",
]

# stage_ms order of the resident evaluation
STAGES = ["centres", "list", "pairs", "pme", "corrections", "pair_kernel", "pme_spread", "pme_fft", "pme_gather"]
TUNING_FIELDS = ["lds_elems", "pair_waves", "spread_tiled", "xcd_remap", "cell_order", "persistent"]


def _check(rc):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        lib().orac_hip_last_error(buf, 1024)
        raise OracHipError(rc, buf.value.decode(errors="replace"))


def _f(a):
    """float64 data in Fortran element order (no copy when it already is)."""
    a = np.asarray(a, dtype=np.float64)
    return np.ascontiguousarray(a) if a.ndim <= 1 else np.asfortranarray(a)


def _i(a):
    a = np.asarray(a, dtype=np.int32)
    return np.ascontiguousarray(a) if a.ndim <= 1 else np.asfortranarray(a)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _out(a, dtype):
    """An in/out argument: must already be a contiguous array of the right dtype."""
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not (a.flags.c_contiguous or a.flags.f_contiguous):
        raise TypeError("in/out arrays must be contiguous numpy arrays of dtype %s" % np.dtype(dtype))
    return a


class Context:
    """One device, one stream, cached topology (orac_hip_ctx)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        _check(lib().orac_hip_create(C.byref(self._h), C.c_int32(device)))

    def close(self):
        if self._h:
            _check(lib().orac_hip_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_name(self):
        buf = C.create_string_buffer(64)
        _check(lib().orac_hip_device_name(self._h, buf, 64))
        return buf.value.decode()

    def probe_copy(self, nbytes):
        """One device copy kernel of nbytes (read) + nbytes (written); calibrates byte counters."""
        _check(lib().orac_hip_probe_copy(self._h, C.c_int64(nbytes)))

    def invalidate(self):
        _check(lib().orac_hip_invalidate(self._h))

    # -- fftw3_pme_init.f90:2
    def fft_pme_init(self, numatoms, nfft1, nfft2, nfft3, order, bsp_mod1=None, bsp_mod2=None, bsp_mod3=None):
        b = [np.zeros(k) if x is None else _out(x, np.float64) for k, x in zip((nfft1, nfft2, nfft3),
                                                                                 (bsp_mod1, bsp_mod2, bsp_mod3))]
        _check(lib().orac_hip_fft_pme_init(self._h, C.c_int32(numatoms), C.c_int32(nfft1), C.c_int32(nfft2),
                                           C.c_int32(nfft3), C.c_int32(order), _p(b[0]), _p(b[1]), _p(b[2])))
        return b

    # -- fft_pme.f:1
    def fft_pme(self, ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, co, oc, volume, alphal, order, nfft1, nfft2, nfft3,
                fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut):
        """Returns eer; fpx/fpy/fpz (+=) and stressc (+=) are updated in place."""
        eer = C.c_double()
        args = [_f(xp), _f(yp), _f(zp), _f(xpcm), _f(ypcm), _f(zpcm), _f(pmechg), _f(co), _f(oc)]
        fpx, fpy, fpz, stressc = (_out(a, np.float64) for a in (fpx, fpy, fpz, stressc))
        ap, gp = _i(atomp), _i(grppt)
        _check(lib().orac_hip_fft_pme(self._h, C.c_int32(ntap), *[_p(a) for a in args], C.c_double(volume),
                                      C.c_double(alphal), C.c_int32(order), C.c_int32(nfft1), C.c_int32(nfft2),
                                      C.c_int32(nfft3), C.byref(eer), _p(fpx), _p(fpy), _p(fpz), _p(stressc),
                                      _p(ap), _p(gp), C.c_int32(1 if pressure else 0), C.c_double(rkcut)))
        return eer.value

    def fft_pme_noalloc(self, ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, co, oc, volume, alphal, order, nfft1,
                        nfft2, nfft3, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut):
        """fft_pme through orac_hip_fft_pme_noalloc (the static scratch region, no allocation)."""
        eer = C.c_double()
        args = [_f(xp), _f(yp), _f(zp), _f(xpcm), _f(ypcm), _f(zpcm), _f(pmechg), _f(co), _f(oc)]
        fpx, fpy, fpz, stressc = (_out(a, np.float64) for a in (fpx, fpy, fpz, stressc))
        ap, gp = _i(atomp), _i(grppt)
        _check(lib().orac_hip_fft_pme_noalloc(
            self._h, C.c_int32(ntap), *[_p(a) for a in args], C.c_double(volume), C.c_double(alphal),
            C.c_int32(order), C.c_int32(nfft1), C.c_int32(nfft2), C.c_int32(nfft3), C.byref(eer), _p(fpx), _p(fpy),
            _p(fpz), _p(stressc), _p(ap), _p(gp), C.c_int32(1 if pressure else 0), C.c_double(rkcut)))
        return eer.value

    # -- fft_pme_alchemy.f:1
    def fft_pme_alchemy(self, ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, charg0, co, oc, volume, alphal, order,
                        nfft1, nfft2, nfft3, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut):
        """Returns (eer, dwrk); fpx/fpy/fpz (+=) and stressc (+=) are updated in place."""
        eer, dw = C.c_double(), C.c_double()
        args = [_f(xp), _f(yp), _f(zp), _f(xpcm), _f(ypcm), _f(zpcm), _f(pmechg), _f(charg0), _f(co), _f(oc)]
        fpx, fpy, fpz, stressc = (_out(a, np.float64) for a in (fpx, fpy, fpz, stressc))
        ap, gp = _i(atomp), _i(grppt)
        _check(lib().orac_hip_fft_pme_alchemy(
            self._h, C.c_int32(ntap), *[_p(a) for a in args], C.c_double(volume), C.c_double(alphal),
            C.c_int32(order), C.c_int32(nfft1), C.c_int32(nfft2), C.c_int32(nfft3), C.byref(eer), _p(fpx), _p(fpy),
            _p(fpz), _p(stressc), _p(ap), _p(gp), C.c_int32(1 if pressure else 0), C.c_double(rkcut), C.byref(dw)))
        return eer.value, dw.value

    # -- lc_index.f:3
    @staticmethod
    def lc_index(indmax, ncx, ncy, ncz, ctoff, co):
        """Returns (nind[2], indxi, indxj, indxk) with indx arrays (2, indmax) flattened."""
        nind = np.zeros(2, np.int32)
        ii, jj, kk = (np.zeros(2 * indmax, np.int32) for _ in range(3))
        _check(lib().orac_hip_lc_index(C.c_int32(indmax), C.c_int32(ncx), C.c_int32(ncy), C.c_int32(ncz), _p(nind),
                                       _p(ii), _p(jj), _p(kk), C.c_double(ctoff), _p(_f(co))))
        return nind, ii, jj, kk

    # -- lc_list.f:1
    def lc_list(self, ncx, ncy, ncz, nind, indxi, indxj, indxk, rcut, co, xp, yp, zp, natp, nnlpp):
        """Fills nnlpp (per group [nv, l...]); returns the number of entries written."""
        nnlpp = _out(nnlpp, np.int32)
        n = C.c_int64()
        _check(lib().orac_hip_lc_list(self._h, C.c_int32(ncx), C.c_int32(ncy), C.c_int32(ncz), _p(_i(nind)),
                                      _p(_i(indxi)), _p(_i(indxj)), _p(_i(indxk)), C.c_double(rcut), _p(_f(co)),
                                      _p(_f(xp)), _p(_f(yp)), _p(_f(zp)), C.c_int32(natp), _p(nnlpp),
                                      C.c_int64(nnlpp.size), C.byref(n)))
        return n.value

    # -- force.f:1
    def force(self, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm, zpcm,
              groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, fpx, fpy, fpz, stress, nnlpp, nnlpp1,
              rneigh, rinn, rout, rtolinn, rtolout, gmass, iz, mass, massinfty, lskip_ewald, ewald, alchemy=None,
              rem=None):
        """Returns a dict with the six energies, ncount1 and the updated rtolinn.

        fpx/fpy/fpz and stress are accumulated in place, nnlpp1 filled.
        rem = (rematom, rem_segkind, rem_factor3, lskip_force, ljintra) calls
        force_rem instead (force_REM.f) and adds seg_ene to the result.
        """
        e = [C.c_double() for _ in range(6)]
        nc1 = C.c_int32()
        rti = C.c_double(rtolinn)
        fpx, fpy, fpz, stress, nnlpp1 = (_out(a, t) for a, t in ((fpx, np.float64), (fpy, np.float64),
                                                                  (fpz, np.float64), (stress, np.float64),
                                                                  (nnlpp1, np.int32)))
        keep = [_i(ss_index), _f(xp0), _f(yp0), _f(zp0), _f(xpg), _f(ypg), _f(zpg), _f(charge), _i(nbtype), _i(type),
                _i(atomg), _f(xpcm), _f(ypcm), _f(zpcm), _i(groupp) if groupp is not None else np.zeros(1, np.int32),
                _i(atomp), _f(co), _f(ecc12), _f(ecc6), _i(mapnl), _i(grppt), _i(nnlpp), _f(gmass), _f(mass)]
        (ss, x, y, z, xg, yg, zg, q, nbt, tt, atg, xc, yc, zc, grp, atp, co_, e12, e6, mp, gpt, nl, gm, ms) = keep
        common = [
            self._h, _p(ss), _p(x), _p(y), _p(z), _p(xg), _p(yg), _p(zg), _p(q), _p(nbt), _p(tt), C.c_int32(ma),
            C.c_int32(nato), _p(atg), _p(xc), _p(yc), _p(zc), _p(grp), _p(atp), _p(co_), _p(e12), _p(e6),
            C.c_double(alphal), _p(mp), C.c_int32(ngrp), _p(gpt), *[C.byref(v) for v in e], _p(fpx), _p(fpy),
            _p(fpz), _p(stress), _p(nl), _p(nnlpp1), C.c_int64(nnlpp1.size), C.c_double(rneigh), C.c_double(rinn),
            C.c_double(rout), C.byref(rti), C.c_double(rtolout), _p(gm), C.c_int32(iz), C.byref(nc1), _p(ms),
            C.c_double(massinfty)]
        flags = [C.c_int32(int(bool(lskip_ewald))), C.c_int32(int(bool(ewald)))]
        dw = None
        seg = None
        if rem is not None:
            rematom, kind, f3, lskip_force, ljintra = rem
            ra = _i(np.asarray(rematom) != 0)
            seg = C.c_double()
            _check(lib().orac_hip_force_rem(*common, C.byref(seg), _p(ra), *flags, C.c_int32(int(kind)),
                                            C.c_double(float(f3)), C.c_int32(int(bool(lskip_force))),
                                            C.c_int32(int(bool(ljintra)))))
        elif alchemy is None:
            _check(lib().orac_hip_force(*common, *flags))
        else:
            al, lam, dwrk = alchemy
            dw = C.c_double(dwrk)
            arrs = [_i(al.ss_alchemy), _f(lam), _f(al.rmin), _f(al.epsm)]
            fn = lib().orac_hip_force_alchemy_shift if al.shifted else lib().orac_hip_force_alchemy
            qx, ljx = al.soft_params()
            _check(fn(
                *common, _p(arrs[0]), _p(arrs[1]), C.c_int32(int(al.intramol)), C.c_int32(int(al.ljintra)),
                C.c_int32(int(al.opls)), _p(arrs[2]), _p(arrs[3]), C.byref(dw), *flags, C.c_double(qx),
                C.c_double(ljx)))
        names = ["uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"]
        out = {k: v.value for k, v in zip(names, e)}
        out.update(ncount1=nc1.value, rtolinn=rti.value)
        if seg is not None:
            out["seg_ene"] = seg.value
        if dw is not None:
            out["dwrk"] = dw.value
        return out

    def force_alchemy(self, *args, alchemy, lam, dwrk=0.0, **kw):
        """force_alchemy.f: the arguments of force(), plus the Alchemy
        description (orac_hip.synth.Alchemy), the eight lambdas (ladd, lrm,
        ladd0, lrm0, laddq, lrmq, laddq0, lrmq0) and the incoming dwrk.  The
        returned dict adds dwrk (incoming value + U(lambda) - U(lambda0))."""
        return self.force(*args, alchemy=(alchemy, lam, dwrk), **kw)

    # -- fptors.f
    def fptors(self, ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, fpx, fpy, fpz, mb):
        """Returns (utors_slt, utors_slv); fpx/fpy/fpz accumulated in place."""
        a, b = C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _i(ltor), _f(xp0), _f(yp0), _f(zp0), _f(pota), _f(potb)]
        _check(lib().orac_hip_fptors(self._h, _p(keep[0]), _p(keep[1]), C.c_int32(ltors), _p(keep[2]), _p(keep[3]),
                                     _p(keep[4]), _p(keep[5]), _p(keep[6]), C.byref(a), C.byref(b), _p(fpx), _p(fpy),
                                     _p(fpz), C.c_int32(mb)))
        return a.value, b.value

    # -- fitors.f
    def fitors(self, ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, fpx, fpy, fpz, mb):
        """Returns (utors_slt, utors_slv); fpx/fpy/fpz accumulated in place."""
        a, b = C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _i(litr), _f(xp0), _f(yp0), _f(zp0), _f(pota), _f(potb), _f(potc)]
        _check(lib().orac_hip_fitors(self._h, _p(keep[0]), _p(keep[1]), C.c_int32(litor), *[_p(v) for v in keep[2:]],
                                     C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), C.c_int32(mb)))
        return a.value, b.value

    # -- fptors_rem.f, fitors_rem.f
    def fptors_rem(self, ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, fpx, fpy, fpz, mb, remltor, rem_factor2):
        """Returns (utors_slt, utors_slv, seg_ene); fpx/fpy/fpz accumulated in place."""
        a, b, g = C.c_double(), C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _i(ltor), _f(xp0), _f(yp0), _f(zp0), _f(pota), _f(potb), _i(np.asarray(remltor) != 0)]
        _check(lib().orac_hip_fptors_rem(self._h, _p(keep[0]), _p(keep[1]), C.c_int32(ltors), *[_p(v) for v in keep[2:7]],
                                         C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _p(keep[7]), C.byref(g),
                                         C.c_double(rem_factor2), C.c_int32(mb)))
        return a.value, b.value, g.value

    def fitors_rem(self, ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, fpx, fpy, fpz, mb, remitor,
                   rem_factor2):
        """Returns (utors_slt, utors_slv, seg_ene); fpx/fpy/fpz accumulated in place."""
        a, b, g = C.c_double(), C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _i(litr), _f(xp0), _f(yp0), _f(zp0), _f(pota), _f(potb), _f(potc),
                _i(np.asarray(remitor) != 0)]
        _check(lib().orac_hip_fitors_rem(self._h, _p(keep[0]), _p(keep[1]), C.c_int32(litor), *[_p(v) for v in keep[2:8]],
                                         C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), _p(keep[8]), C.byref(g),
                                         C.c_double(rem_factor2), C.c_int32(mb)))
        return a.value, b.value, g.value

    # -- fpbond.f
    def fpbond(self, ss_index, lbnd, lbond, xp0, yp0, zp0, pota, potb, fpx, fpy, fpz, mb):
        """Returns (ubond_slt, ubond_slv); fpx/fpy/fpz accumulated in place."""
        a, b = C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _i(lbnd), _f(xp0), _f(yp0), _f(zp0), _f(pota), _f(potb)]
        _check(lib().orac_hip_fpbond(self._h, _p(keep[0]), _p(keep[1]), C.c_int32(lbond), _p(keep[2]), _p(keep[3]),
                                     _p(keep[4]), _p(keep[5]), _p(keep[6]), C.byref(a), C.byref(b), _p(fpx), _p(fpy),
                                     _p(fpz), C.c_int32(mb)))
        return a.value, b.value

    # -- fpbend.f
    def fpbend(self, ss_index, lbndg, lbend, xp0, yp0, zp0, pota, potb, potc, potd, pote, potf, potg, fpx, fpy, fpz,
               mb):
        """Returns (ubend_slt, ubend_slv); fpx/fpy/fpz accumulated in place."""
        a, b = C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _i(lbndg), _f(xp0), _f(yp0), _f(zp0)] + [_f(v) for v in
                                                                       (pota, potb, potc, potd, pote, potf, potg)]
        _check(lib().orac_hip_fpbend(self._h, _p(keep[0]), _p(keep[1]), C.c_int32(lbend), *[_p(v) for v in keep[2:]],
                                     C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz), C.c_int32(mb)))
        return a.value, b.value

    # -- fnb14.f
    def fnb14(self, ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, int14, int14p, type14,
              fudge, lj_fudge, fpx, fpy, fpz):
        """Returns (uconf_slt, uconf_slv, ucoul_slt, ucoul_slv); forces accumulated in place."""
        e = [C.c_double() for _ in range(4)]
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _f(xp0), _f(yp0), _f(zp0), _f(charge), _f(ecc12), _f(ecc6), _i(int14), _i(type14)]
        _check(lib().orac_hip_fnb14(self._h, _p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]),
                                    C.c_int32(nato), _p(keep[5]), _p(keep[6]), C.c_int32(int(bool(cut))),
                                    C.c_int32(int(bool(ewald))), C.c_double(alphal), _p(keep[7]), C.c_int32(int14p),
                                    _p(keep[8]), C.c_double(fudge), C.c_double(lj_fudge), _p(fpx), _p(fpy), _p(fpz),
                                    *[C.byref(v) for v in e]))
        return tuple(v.value for v in e)

    # -- fnb14_rem.f
    def fnb14_rem(self, ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, int14, int14p, type14,
                  fudge, lj_fudge, fpx, fpy, fpz, remint14, rem_factor2, seg_ene=0.0):
        """Returns (uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, seg_ene); forces accumulated in place.
        seg_ene is the incoming value, returned unchanged when there is no pair."""
        e = [C.c_double() for _ in range(4)]
        g = C.c_double(seg_ene)
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _f(xp0), _f(yp0), _f(zp0), _f(charge), _f(ecc12), _f(ecc6), _i(int14), _i(type14),
                _i(np.asarray(remint14) != 0)]
        _check(lib().orac_hip_fnb14_rem(self._h, _p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]),
                                        C.c_int32(nato), _p(keep[5]), _p(keep[6]), C.c_int32(int(bool(cut))),
                                        C.c_int32(int(bool(ewald))), C.c_double(alphal), _p(keep[7]), C.c_int32(int14p),
                                        _p(keep[8]), C.c_double(fudge), C.c_double(lj_fudge), _p(fpx), _p(fpy),
                                        _p(fpz), *[C.byref(v) for v in e], _p(keep[9]), C.byref(g),
                                        C.c_double(rem_factor2)))
        return tuple(v.value for v in e) + (g.value,)

    # -- mts_forces.f:212
    @staticmethod
    def decide_force(rinn, alphal):
        v = C.c_int32()
        _check(lib().orac_hip_decide_force(C.c_double(rinn), C.c_double(alphal), C.byref(v)))
        return bool(v.value)

    # -- ferrf.f:1
    def ferrf(self, ss_index, alphal, charge, fudge, x0, y0, z0, iz, list, nlist, fpx, fpy, fpz, erf_corr=False,
              nbonds_added=0, mb=None, erf_arr_corr=None, delew=0.0, rlew=0.0):
        """Returns (fscnstr_slt, fscnstr_slv); forces accumulated in place when iz = 1.
        erf_corr uses the ERF_CORR table erf_arr_corr (4, n) with spacing delew from rlew."""
        a, b = C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        mb = len(fpx) if mb is None else mb
        keep = [_i(ss_index), _f(charge), _f(x0), _f(y0), _f(z0), _i(list),
                _f(erf_arr_corr) if erf_arr_corr is not None else np.zeros(4)]
        _check(lib().orac_hip_ferrf(self._h, _p(keep[0]), C.c_double(alphal), _p(keep[1]), C.c_double(fudge),
                                    _p(keep[2]), _p(keep[3]), _p(keep[4]), C.c_int32(iz), _p(keep[5]),
                                    C.c_int32(nlist), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz),
                                    C.c_int32(int(bool(erf_corr))), _p(keep[6]), C.c_double(delew),
                                    C.c_double(rlew), C.c_int32(nbonds_added), C.c_int32(mb)))
        return a.value, b.value

    @staticmethod
    def int_corr_erf_spline(rmin, rmax, nbin, alpha, kcut):
        """The ERF_CORR spline table: returns (erf_arr_corr (4, nbin + 1) Fortran order, tau)."""
        tab = np.zeros((4, nbin + 1), order="F")
        tau = np.zeros(nbin + 1)
        _check(lib().orac_hip_int_corr_erf_spline(C.c_double(rmin), C.c_double(rmax), C.c_int32(nbin),
                                                  C.c_double(alpha), C.c_double(kcut), _p(tab), _p(tau)))
        return tab, tau

    @staticmethod
    def erf_corr_cutoff(oc, nfft1, nfft2, nfft3):
        rk, dl = C.c_double(), C.c_double()
        _check(lib().orac_hip_erf_corr_cutoff(_p(_f(oc)), C.byref(dl), C.byref(rk), C.c_int32(nfft1),
                                              C.c_int32(nfft2), C.c_int32(nfft3)))
        return rk.value

    # -- ferrf_alchemy.f:1
    def ferrf_alchemy(self, ss_index, alphal, charge, charge0, fudge, x0, y0, z0, iz, list, nlist, fpx, fpy, fpz,
                      nbonds_added=0, mb=None):
        """Returns (fscnstr_slt, fscnstr_slv, dwrk); forces accumulated in place when iz = 1."""
        a, b, dw = C.c_double(), C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        mb = len(fpx) if mb is None else mb
        keep = [_i(ss_index), _f(charge), _f(charge0), _f(x0), _f(y0), _f(z0), _i(list)]
        _check(lib().orac_hip_ferrf_alchemy(self._h, _p(keep[0]), C.c_double(alphal), _p(keep[1]), _p(keep[2]),
                                            C.c_double(fudge), _p(keep[3]), _p(keep[4]), _p(keep[5]), C.c_int32(iz),
                                            _p(keep[6]), C.c_int32(nlist), C.byref(a), C.byref(b), _p(fpx), _p(fpy),
                                            _p(fpz), C.c_int32(0), C.c_int32(nbonds_added), C.byref(dw),
                                            C.c_int32(mb)))
        return a.value, b.value, dw.value

    # -- cself_alchemy.f:1
    def cself_alchemy(self, ss_index, nato, alphal, kcut, charge, charge0):
        """Returns (self_slt, dwrk)."""
        a, dw = C.c_double(), C.c_double()
        _check(lib().orac_hip_cself_alchemy(self._h, _p(_i(ss_index)), C.c_int32(nato), C.c_double(alphal),
                                            C.c_double(kcut), _p(_f(charge)), _p(_f(charge0)), C.byref(a),
                                            C.byref(dw)))
        return a.value, dw.value

    # -- cself.f:1
    def cself(self, ss_index, nato, alphal, kcut, charge):
        a, b = C.c_double(), C.c_double()
        _check(lib().orac_hip_cself(self._h, _p(_i(ss_index)), C.c_int32(nato), C.c_double(alphal), C.c_double(kcut),
                                    _p(_f(charge)), C.byref(a), C.byref(b)))
        return a.value, b.value

    # -- fnbgrp.f:1
    def fnbgrp(self, ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, ewald, alphal, ingrp, ingrpp, fpx, fpy, fpz,
               nato=None):
        """Returns (uconf_slt, uconf_slv, ucoul_slt, ucoul_slv); forces accumulated in place."""
        e = [C.c_double() for _ in range(4)]
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        nato = len(fpx) if nato is None else nato
        keep = [_i(ss_index), _f(xp0), _f(yp0), _f(zp0), _f(charge), _i(type), _f(ecc12), _f(ecc6), _i(ingrp)]
        _check(lib().orac_hip_fnbgrp(self._h, *[_p(k) for k in keep[:8]], C.c_int32(int(bool(ewald))),
                                     C.c_double(alphal), _p(keep[8]), C.c_int32(ingrpp), _p(fpx), _p(fpy), _p(fpz),
                                     *[C.byref(v) for v in e], C.c_int32(nato)))
        return tuple(v.value for v in e)

    # -- ferrf_alchemy_add.f:1
    def ferrf_alchemy_add(self, ss_index, alphal, charge, lambdaq, lambdaq0, x0, y0, z0, list, nlist, fpx, fpy, fpz,
                          erf_corr=False, erf_arr_corr=None, delew=0.0, rlew=0.0, nato=None, intramol=False):
        """Returns (fscnstr_slt, fscnstr_slv, dwrk); forces accumulated in place.
        list is (3, nlist): ia, ib (1-based) and facQ."""
        a, b, dw = C.c_double(), C.c_double(), C.c_double()
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        nato = len(fpx) if nato is None else nato
        keep = [_i(ss_index), _f(charge), _f(lambdaq), _f(lambdaq0), _f(x0), _f(y0), _f(z0), _i(list),
                _f(erf_arr_corr) if erf_arr_corr is not None else np.zeros(4)]
        _check(lib().orac_hip_ferrf_alchemy_add(self._h, _p(keep[0]), C.c_double(alphal), *[_p(k) for k in keep[1:8]],
                                                C.c_int32(nlist), C.byref(a), C.byref(b), _p(fpx), _p(fpy), _p(fpz),
                                                C.c_int32(int(bool(erf_corr))), _p(keep[8]), C.c_double(delew),
                                                C.c_double(rlew), C.byref(dw), C.c_int32(nato),
                                                C.c_int32(int(bool(intramol)))))
        return a.value, b.value, dw.value

    # -- fnb14_alchemy.f:1
    def fnb14_alchemy(self, ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, int14, int14p,
                      type14, fudge, lj_fudge, fpx, fpy, fpz, lambda_, lambdaq, lambda0, lambdaq0):
        """Returns (uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, dwrk); forces accumulated in place."""
        e = [C.c_double() for _ in range(5)]
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        keep = [_i(ss_index), _f(xp0), _f(yp0), _f(zp0), _f(charge), _f(ecc12), _f(ecc6), _i(int14), _i(type14),
                _f(lambda_), _f(lambdaq), _f(lambda0), _f(lambdaq0)]
        _check(lib().orac_hip_fnb14_alchemy(self._h, _p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]),
                                            C.c_int32(nato), _p(keep[5]), _p(keep[6]), C.c_int32(int(bool(cut))),
                                            C.c_int32(int(bool(ewald))), C.c_double(alphal), _p(keep[7]),
                                            C.c_int32(int14p), _p(keep[8]), C.c_double(fudge), C.c_double(lj_fudge),
                                            _p(fpx), _p(fpy), _p(fpz), *[C.byref(v) for v in e[:4]],
                                            *[_p(k) for k in keep[9:]], C.byref(e[4])))
        return tuple(v.value for v in e)

    # -- fnbgrp_alchemy.f:1
    def fnbgrp_alchemy(self, ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, ewald, alphal, ingrp, ingrpp, fpx,
                       fpy, fpz, lambdaq, nato=None):
        """Returns (uconf_slt, uconf_slv, ucoul_slt, ucoul_slv); forces accumulated in place."""
        e = [C.c_double() for _ in range(4)]
        fpx, fpy, fpz = (_out(v, np.float64) for v in (fpx, fpy, fpz))
        nato = len(fpx) if nato is None else nato
        keep = [_i(ss_index), _f(xp0), _f(yp0), _f(zp0), _f(charge), _i(type), _f(ecc12), _f(ecc6), _i(ingrp),
                _f(lambdaq)]
        _check(lib().orac_hip_fnbgrp_alchemy(self._h, *[_p(k) for k in keep[:8]], C.c_int32(int(bool(ewald))),
                                             C.c_double(alphal), _p(keep[8]), C.c_int32(ingrpp), _p(fpx), _p(fpy),
                                             _p(fpz), *[C.byref(v) for v in e], _p(keep[9]), C.c_int32(nato)))
        return tuple(v.value for v in e)

    def sync(self):
        _check(lib().orac_hip_sync(self._h))


class _SysDesc(C.Structure):
    _fields_ = [
        ("natoms", C.c_int32), ("ngrp", C.c_int32), ("nmol", C.c_int32), ("ntypes", C.c_int32),
        ("charge", C.c_void_p), ("mass", C.c_void_p), ("nbtype", C.c_void_p), ("type_table", C.c_void_p),
        ("ecc12", C.c_void_p), ("ecc6", C.c_void_p), ("grppt", C.c_void_p), ("atomp", C.c_void_p),
        ("ss_index", C.c_void_p), ("mapnl", C.c_void_p), ("ncorr", C.c_int32), ("corr_pairs", C.c_void_p),
        ("corr_fudge", C.c_void_p), ("co", C.c_void_p), ("alpha", C.c_double), ("nfft1", C.c_int32),
        ("nfft2", C.c_int32), ("nfft3", C.c_int32), ("order", C.c_int32), ("rcut", C.c_double), ("rtol", C.c_double),
        ("rnei", C.c_double), ("ncx", C.c_int32), ("ncy", C.c_int32), ("ncz", C.c_int32), ("precision", C.c_int32),
    ]


ENERGY_FIELDS = ["eer", "uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss", "corr_slt",
                 "corr_slv"]


class _AlchDesc(C.Structure):
    """orac_hip_alchemy_desc (include/orac_hip.h)."""
    _fields_ = [("ss_alchemy", C.c_void_p), ("rmin", C.c_void_p), ("epsm", C.c_void_p), ("qsoft", C.c_double),
                ("ljsoft", C.c_double), ("intramol", C.c_int32), ("ljintra", C.c_int32), ("opls", C.c_int32),
                ("shifted", C.c_int32), ("qshift", C.c_double), ("ljshift", C.c_double), ("nqq", C.c_int32),
                ("listqq", C.c_void_p)]


class ResidentSystem:
    """A system kept in HBM: topology once, coordinates per step, forces on the device.

    ``s`` is any object with the arrays of ``orac_hip.synth.System``.
    ``corr_pairs``/``corr_fudge`` are the per-step reciprocal corrections
    (ferrf with iz = 1, e.g. flexible bonds and bends).
    """

    def __init__(self, ctx, s, *, precision=0, ncell=None, corr_pairs=None, corr_fudge=None):
        self.ctx = ctx
        self.natoms = s.natoms
        nc = ncell or max(3, int(s.L // 3.5))
        mapnl = s.mapnl()
        cp = np.zeros((2, 0), np.int32) if corr_pairs is None else np.asarray(corr_pairs, np.int32)
        cf = np.ones(cp.shape[1]) if corr_fudge is None else np.asarray(corr_fudge, np.float64)
        self._keep = dict(charge=_f(s.charge), mass=_f(s.mass), nbtype=_i(s.nbtype), type_table=_i(s.type_table),
                          ecc12=_f(s.ecc12), ecc6=_f(s.ecc6), grppt=_i(s.grppt), atomp=_i(s.atomp),
                          ss_index=_i(s.ss_index), mapnl=_i(mapnl), corr_pairs=_i(cp), corr_fudge=_f(cf),
                          co=_f(s.co))
        k = self._keep
        d = _SysDesc(natoms=s.natoms, ngrp=s.ngrp, nmol=s.nprot, ntypes=s.ntypes, ncorr=cp.shape[1],
                     alpha=s.alpha, nfft1=s.nfft[0], nfft2=s.nfft[1], nfft3=s.nfft[2], order=s.order, rcut=s.rcut,
                     rtol=s.rtol, rnei=s.rnei, ncx=nc, ncy=nc, ncz=nc, precision=precision)
        for name in ("charge", "mass", "nbtype", "type_table", "ecc12", "ecc6", "grppt", "atomp", "ss_index", "mapnl",
                     "corr_pairs", "corr_fudge", "co"):
            setattr(d, name, k[name].ctypes.data)
        self._desc = d
        self.id = C.c_int32()
        _check(lib().orac_hip_system_create(ctx._h, C.byref(d), C.byref(self.id)))
        self.set_coords(s.xp0, s.yp0, s.zp0)

    def set_coords(self, xp, yp, zp):
        a, b, c = _f(xp), _f(yp), _f(zp)
        _check(lib().orac_hip_system_set_coords(self.ctx._h, self.id, _p(a), _p(b), _p(c)))

    def eval(self, rebuild_list=True):
        _check(lib().orac_hip_system_eval(self.ctx._h, self.id, C.c_int32(1 if rebuild_list else 0)))

    def results(self):
        n = self.natoms
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        e = np.zeros(21)
        _check(lib().orac_hip_system_results(self.ctx._h, self.id, _p(fx), _p(fy), _p(fz), _p(e)))
        out = {k: e[i] for i, k in enumerate(ENERGY_FIELDS)}
        out.update(f=np.stack([fx, fy, fz]), stress=e[9:18].reshape(3, 3, order="F"), list_entries=int(e[18]),
                   self_slt=e[19], self_slv=e[20])
        return out

    def set_shells(self, rcutm, rtolm):
        """Two nonbonded shells (inner 0 .. rcutm, outer rcutm .. rcut, switched
        over rtolm); rcutm <= 0 goes back to one shell."""
        _check(lib().orac_hip_system_set_shells(self.ctx._h, self.id, C.c_double(rcutm), C.c_double(rtolm)))

    def set_alchemy(self, al, lam=None):
        """Make the system alchemical with orac_hip.synth.Alchemy al (its
        listqq, if set, adds ferrf_alchemy_add's intra-alchemical term), then
        set the eight lambdas lam (synth.lambda_state order) if given."""
        qq = np.zeros((3, 0), np.int32) if al.listqq is None else np.asarray(al.listqq, np.int32)
        self._alch = dict(ss=_i(al.ss_alchemy), rmin=_f(al.rmin), epsm=_f(al.epsm), qq=_i(qq))
        d = _AlchDesc(ss_alchemy=self._alch["ss"].ctypes.data, rmin=self._alch["rmin"].ctypes.data,
                      epsm=self._alch["epsm"].ctypes.data, qsoft=al.qsoft, ljsoft=al.ljsoft,
                      intramol=int(al.intramol), ljintra=int(al.ljintra), opls=int(al.opls),
                      shifted=int(al.shifted), qshift=al.qshift, ljshift=al.ljshift, nqq=qq.shape[1],
                      listqq=self._alch["qq"].ctypes.data)
        _check(lib().orac_hip_system_set_alchemy(self.ctx._h, self.id, C.byref(d)))
        if lam is not None:
            self.set_lambda(lam)

    def set_lambda(self, lam):
        lam = _f(lam)
        assert lam.size == 8
        _check(lib().orac_hip_system_set_lambda(self.ctx._h, self.id, _p(lam)))

    def alchemy_results(self):
        """dwrk = U(lambda) - U(lambda0) of the last eval and its parts."""
        o = np.zeros(6)
        _check(lib().orac_hip_system_alchemy_results(self.ctx._h, self.id, _p(o)))
        return dict(zip(("dwrk", "direct", "recip", "corr", "self", "eer0"), o.tolist()))

    def profile(self, enable=True):
        _check(lib().orac_hip_system_profile(self.ctx._h, self.id, C.c_int32(1 if enable else 0)))

    def stage_ms(self):
        """Milliseconds of the last eval's stages, keyed by ``STAGES``."""
        ms = (C.c_float * len(STAGES))()
        _check(lib().orac_hip_system_stage_ms(self.ctx._h, self.id, ms, C.c_int32(len(STAGES))))
        return dict(zip(STAGES, ms))

    def device_bytes(self):
        b = C.c_int64()
        _check(lib().orac_hip_system_bytes(self.ctx._h, self.id, C.byref(b)))
        return b.value

    def close(self):
        if self.id.value:
            _check(lib().orac_hip_system_destroy(self.ctx._h, self.id))
            self.id = C.c_int32(0)


class Batch:
    """Resident systems evaluated together, one kernel launch per stage."""

    def __init__(self, ctx, systems):
        self.ctx = ctx
        self.systems = list(systems)
        ids = np.array([s.id.value for s in self.systems], np.int32)
        self.id = C.c_int32()
        _check(lib().orac_hip_batch_create(ctx._h, C.c_int32(len(ids)), _p(ids), C.byref(self.id)))

    def eval(self, rebuild_list=True, graph=False):
        _check(lib().orac_hip_batch_eval(self.ctx._h, self.id, C.c_int32(1 if rebuild_list else 0),
                                         C.c_int32(1 if graph else 0)))

    def results(self):
        return [s.results() for s in self.systems]

    def profile(self, enable=True):
        _check(lib().orac_hip_batch_profile(self.ctx._h, self.id, C.c_int32(1 if enable else 0)))

    def stage_ms(self):
        ms = (C.c_float * len(STAGES))()
        _check(lib().orac_hip_batch_stage_ms(self.ctx._h, self.id, ms, C.c_int32(len(STAGES))))
        return dict(zip(STAGES, ms))

    def tuning(self):
        t = np.zeros(len(TUNING_FIELDS), np.int32)
        _check(lib().orac_hip_batch_get_tuning(self.ctx._h, self.id, _p(t), C.c_int32(len(t))))
        return dict(zip(TUNING_FIELDS, (int(v) for v in t)))

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = np.array([cur[k] for k in TUNING_FIELDS], np.int32)
        _check(lib().orac_hip_batch_set_tuning(self.ctx._h, self.id, _p(t), C.c_int32(len(t))))

    def tune_log(self):
        """Every candidate the last autotune timed: a list of (tuning dict, ms per step)."""
        n = C.c_int32()
        _check(lib().orac_hip_batch_tune_log(self.ctx._h, self.id, None, None, C.c_int32(0), C.byref(n)))
        t = np.zeros((max(1, n.value), len(TUNING_FIELDS)), np.int32)
        ms = np.zeros(max(1, n.value))
        _check(lib().orac_hip_batch_tune_log(self.ctx._h, self.id, _p(t), _p(ms), C.c_int32(n.value), C.byref(n)))
        return [(dict(zip(TUNING_FIELDS, (int(v) for v in t[k]))), float(ms[k])) for k in range(n.value)]

    def autotune(self, reps=5, rebuild_list=True):
        """Pick the fastest tuning by timing the batch; returns (tuning, ms per eval)."""
        ms = C.c_double()
        _check(lib().orac_hip_batch_autotune(self.ctx._h, self.id, C.c_int32(reps),
                                             C.c_int32(1 if rebuild_list else 0), C.byref(ms)))
        return self.tuning(), ms.value

    def close(self):
        if self.id.value:
            _check(lib().orac_hip_batch_destroy(self.ctx._h, self.id))
            self.id = C.c_int32(0)
