"""GPU parity: every C-ABI entry point against the CPU oracle on the same inputs.

Tolerances (FP64 plain path): lists bit-exact; forces relative RMS <= 1e-9;
energies and virial <= 1e-10 relative.  The PME grid is accumulated in 64-bit
fixed point, so the GPU's reciprocal sums are deterministic but not the
oracle's bit pattern.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel, relrms

pytestmark = pytest.mark.gpu

F_TOL = 1e-9
E_TOL = 1e-10


def _stress_close(a, b, tol=E_TOL):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


def _pme_gpu(ctx, s, pressure=True):
    n = s.natoms
    ctx.fft_pme_init(n, *s.nfft, s.order)
    cm = s.molecule_coms()
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    st = np.zeros((3, 3), order="F")
    eer = ctx.fft_pme(n, s.xp0, s.yp0, s.zp0, cm[0], cm[1], cm[2], s.charge, s.co, s.oc, s.volume, s.alpha, s.order,
                      *s.nfft, fpx, fpy, fpz, st, s.atomp, s.grppt, pressure, 1e9)
    return dict(eer=eer, f=np.stack([fpx, fpy, fpz]), stress=st)


def _force_gpu(ctx, s, nnlpp, fa, *, rinn, rout, rtolinn, rtolout, rneigh, iz=1, lskip_ewald=False, ewald=True,
               massinfty=1e20, cap=None):
    n = s.natoms
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    st = np.zeros((3, 3), order="F")
    cap = cap or (len(nnlpp) + s.ngrp + 16 if iz == 1 else s.ngrp * (s.ngrp // 2 + 2))
    nn1 = np.zeros(cap, np.int32)
    r = ctx.force(s.ss_index, s.xp0, s.yp0, s.zp0, fa.xpg, fa.ypg, fa.zpg, s.charge, s.nbtype, s.type_table,
                  s.ntypes, n, s.atomg, fa.xpcm, fa.ypcm, fa.zpcm, None, s.atomp, s.co, s.ecc12, s.ecc6, s.alpha,
                  fa.mapnl, s.ngrp, s.grppt, fpx, fpy, fpz, st, nnlpp, nn1, rneigh, rinn, rout, rtolinn, rtolout,
                  s.pmass, iz, s.mass, massinfty, lskip_ewald, ewald)
    r.update(f=np.stack([fpx, fpy, fpz]), stress=st, nnlpp1=nn1[:r["ncount1"]])
    return r


def _oracle_list(s):
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    if s.L > 3 * rc:
        nc = int(s.L // 3.0)
        st = O.lc_index(nc, nc, nc, rc, s.co)
        return O.lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg), fa, (nc, st)
    u = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0, fa=fa)
    return u["nnlpp1"], fa, None


@pytest.fixture(scope="module")
def solvated24k():
    from orac_hip import synth
    return synth.config("solvated24k")


def test_device_is_gfx950(gpu_ctx):
    assert gpu_ctx.device_name().startswith("gfx950")


@pytest.mark.parametrize("order", [4, 5, 6])
@pytest.mark.parametrize("k", [20, 24, 48, 64, 30])
def test_bsp_moduli_exact(gpu_ctx, order, k):
    b = gpu_ctx.fft_pme_init(100, k, k, k, order)
    assert np.array_equal(b[0], O.bsp_moduli(order, k))


@pytest.mark.parametrize("sysname", ["water216", "small_solute", "solvated24k"])
def test_fft_pme(gpu_ctx, sysname, request):
    s = request.getfixturevalue(sysname)
    g = _pme_gpu(gpu_ctx, s)
    o = O.fft_pme(s, variant="fftw")
    print(sysname, "eer rel", rel(g["eer"], o["eer"]), "f relrms", relrms(g["f"], o["f"]))
    assert rel(g["eer"], o["eer"]) < E_TOL
    assert relrms(g["f"], o["f"]) < F_TOL
    assert _stress_close(g["stress"], o["stress"])


@pytest.mark.parametrize("nfft,order", [((30, 24, 20), 4), ((27, 24, 20), 4), ((24, 30, 25), 6)])
def test_fft_pme_noncubic_grid(gpu_ctx, small_solute, nfft, order):
    """Tiled spread (every axis has a tile edge dividing K) and the atomic
    fallback (K = 27 has none in [order-1, 8])."""
    import dataclasses
    s = dataclasses.replace(small_solute, nfft=nfft, order=order)
    g = _pme_gpu(gpu_ctx, s)
    o = O.fft_pme(s, variant="fftw")
    assert rel(g["eer"], o["eer"]) < E_TOL
    assert relrms(g["f"], o["f"]) < F_TOL
    assert _stress_close(g["stress"], o["stress"])


def test_fft_pme_deterministic(gpu_ctx, small_solute):
    a = _pme_gpu(gpu_ctx, small_solute)
    b = _pme_gpu(gpu_ctx, small_solute)
    assert a["eer"] == b["eer"] and np.array_equal(a["f"], b["f"])


def test_fft_pme_accumulates(gpu_ctx, water216):
    s = water216
    n = s.natoms
    gpu_ctx.fft_pme_init(n, *s.nfft, s.order)
    cm = s.molecule_coms()
    f0 = np.random.default_rng(1).normal(size=(3, n))
    fpx, fpy, fpz = f0[0].copy(), f0[1].copy(), f0[2].copy()
    st0 = np.asfortranarray(np.arange(9.0).reshape(3, 3))
    st = st0.copy(order="F")
    gpu_ctx.fft_pme(n, s.xp0, s.yp0, s.zp0, cm[0], cm[1], cm[2], s.charge, s.co, s.oc, s.volume, s.alpha, s.order,
                    *s.nfft, fpx, fpy, fpz, st, s.atomp, s.grppt, True, 1e9)
    g = _pme_gpu(gpu_ctx, s)
    assert np.allclose(np.stack([fpx, fpy, fpz]) - f0, g["f"], rtol=0, atol=1e-12)
    assert np.allclose(st - st0, g["stress"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("sysname", ["small_solute", "solvated24k"])
def test_lc_list_exact(gpu_ctx, sysname, request):
    s = request.getfixturevalue(sysname)
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nc = int(s.L // 3.0)
    nind, ii, jj, kk = gpu_ctx.lc_index(2500, nc, nc, nc, rc, s.co)
    so = O.lc_index(nc, nc, nc, rc, s.co)
    assert np.array_equal(nind, so[0]) and np.array_equal(ii, so[1]) and np.array_equal(jj, so[2])
    ref = O.lc_list(nc, nc, nc, so, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
    out = np.zeros(len(ref) + 100, np.int32)
    n = gpu_ctx.lc_list(nc, nc, nc, nind, ii, jj, kk, rc, s.co, fa.xpg, fa.ypg, fa.zpg, s.ngrp, out)
    assert n == len(ref)
    assert np.array_equal(out[:n], ref)


@pytest.mark.parametrize("sysname", ["small_solute", "solvated24k"])
def test_lc_list_exact_coarse_cells(gpu_ctx, sysname, request):
    """3 cells per side: the stencil has fewer than 64 cells and each cell
    many groups, so the walk takes the members of one cell across the lanes;
    the list must still be the reference's, entry for entry."""
    s = request.getfixturevalue(sysname)
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nc = 3
    so = O.lc_index(nc, nc, nc, rc, s.co)
    assert so[0][1] < 64
    nind, ii, jj, kk = gpu_ctx.lc_index(2500, nc, nc, nc, rc, s.co)
    ref = O.lc_list(nc, nc, nc, so, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
    out = np.zeros(len(ref) + 100, np.int32)
    n = gpu_ctx.lc_list(nc, nc, nc, nind, ii, jj, kk, rc, s.co, fa.xpg, fa.ypg, fa.zpg, s.ngrp, out)
    assert n == len(ref)
    assert np.array_equal(out[:n], ref)


def test_lc_list_capacity_error(gpu_ctx, small_solute):
    from orac_hip.api import OracHipError
    s = small_solute
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nc = int(s.L // 3.0)
    nind, ii, jj, kk = gpu_ctx.lc_index(2500, nc, nc, nc, rc, s.co)
    with pytest.raises(OracHipError) as e:
        gpu_ctx.lc_list(nc, nc, nc, nind, ii, jj, kk, rc, s.co, fa.xpg, fa.ypg, fa.zpg, s.ngrp,
                        np.zeros(10, np.int32))
    assert e.value.code == 2


SHELLS = [
    dict(rinn=0.0, rout=None, rtolinn=0.0, rtolout=0.3, rneigh=0.0),
    dict(rinn=4.0, rout=7.0, rtolinn=0.5, rtolout=0.5, rneigh=0.6),
    dict(rinn=2.5, rout=5.1, rtolinn=0.3, rtolout=0.3, rneigh=0.4),
]


@pytest.mark.parametrize("shell", range(len(SHELLS)))
@pytest.mark.parametrize("sysname", ["water216", "small_solute", "solvated24k"])
def test_force(gpu_ctx, sysname, shell, request):
    s = request.getfixturevalue(sysname)
    nnlpp, fa, _ = _oracle_list(s)
    sh = dict(SHELLS[shell])
    if sh["rout"] is None:
        sh["rout"] = s.rcut
    g = _force_gpu(gpu_ctx, s, nnlpp, fa, **sh)
    o = O.force(s, nnlpp, fa=fa, **sh)
    print(sysname, shell, "f relrms", relrms(g["f"], o["f"]))
    assert relrms(g["f"], o["f"]) < F_TOL
    for k in ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"):
        if o[k] == 0.0:
            assert g[k] == 0.0, k
        else:
            assert rel(g[k], o[k]) < E_TOL, (k, g[k], o[k])
    assert _stress_close(g["stress"], o["stress"])
    assert g["ncount1"] == o["ncount1"] and g["rtolinn"] == o["rtolinn"]
    assert np.array_equal(g["nnlpp1"], o["nnlpp1"])


@pytest.mark.parametrize("lskip,ewald", [(True, True), (True, False)])
def test_force_lskip_plain_coulomb(gpu_ctx, small_solute, lskip, ewald):
    s = small_solute
    nnlpp, fa, _ = _oracle_list(s)
    kw = dict(rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0, lskip_ewald=lskip, ewald=ewald)
    g = _force_gpu(gpu_ctx, s, nnlpp, fa, **kw)
    o = O.force(s, nnlpp, fa=fa, **kw)
    assert relrms(g["f"], o["f"]) < F_TOL
    assert rel(g["uconf_ss"], o["uconf_ss"]) < E_TOL
    if o["ucoul_slv"]:
        assert rel(g["ucoul_slv"], o["ucoul_slv"]) < E_TOL


def test_force_massinfty_rule(gpu_ctx, small_solute):
    """Atoms at or above massinfty skip masked pairs only (force.f:488-489)."""
    import dataclasses
    s = small_solute
    mass = s.mass.copy()
    mass[:20] = 1e21
    s2 = dataclasses.replace(s, mass=mass)
    nnlpp, fa, _ = _oracle_list(s2)
    kw = dict(rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0, massinfty=1e20)
    g = _force_gpu(gpu_ctx, s2, nnlpp, fa, **kw)
    o = O.force(s2, nnlpp, fa=fa, **kw)
    assert relrms(g["f"], o["f"]) < F_TOL
    assert rel(g["uconf_slt"], o["uconf_slt"]) < E_TOL


@pytest.mark.parametrize("sysname", ["water216", "small_solute"])
def test_force_phony_list_call(gpu_ctx, sysname, request):
    """iz = 0: the O(N^2) list update over [ngrp, 1..ngrp] (mts_forces 'u' shell)."""
    s = request.getfixturevalue(sysname)
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    allg = O.all_groups_list(s.ngrp)
    kw = dict(rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0)
    g = _force_gpu(gpu_ctx, s, allg, fa, **kw)
    o = O.force(s, allg, fa=fa, **kw)
    assert g["ncount1"] == o["ncount1"]
    assert np.array_equal(g["nnlpp1"], o["nnlpp1"])
    assert not g["f"].any()


@pytest.mark.parametrize("which,iz,fudge", [("constraints", 0, 1.0), ("bonds", 1, 1.0), ("bends", 1, 0.8333)])
def test_ferrf(gpu_ctx, small_solute, which, iz, fudge):
    s = small_solute
    pairs = getattr(s, which)
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.ferrf(s.ss_index, s.alpha, s.charge, fudge, x, y, z, iz, pairs, pairs.shape[1], fpx, fpy, fpz)
    o = O.ferrf(s, pairs, iz, fudge)
    for g_, o_ in ((a, o["slt"]), (b, o["slv"])):
        assert (g_ == 0.0 and o_ == 0.0) or rel(g_, o_) < E_TOL
    if iz == 1:
        assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < F_TOL
    else:
        assert not fpx.any()


def test_cself(gpu_ctx, small_solute):
    s = small_solute
    for kcut in (1e9, 3.0):
        g = gpu_ctx.cself(s.ss_index, s.natoms, s.alpha, kcut, s.charge)
        o = O.cself(s, kcut)
        for a, b in zip(g, o):
            assert rel(a, b) < 1e-13


def test_fnbgrp(gpu_ctx, water216):
    """Intra-group pairs of a few waters, with and without Ewald."""
    s = water216
    n = s.natoms
    ingrp = np.array([[1, 2], [4, 6], [10, 11]], np.int32).T.copy(order="F")
    # per-atom LJ types: O = 1, H = 2 (packed index j(j-1)/2 + i)
    typ = np.where(np.arange(n) % 3 == 0, 1, 2).astype(np.int32)
    for ewald in (True, False):
        fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
        g = gpu_ctx.fnbgrp(s.ss_index, s.xp0, s.yp0, s.zp0, s.charge, typ, s.ecc12, s.ecc6, ewald, s.alpha, ingrp,
                           ingrp.shape[1], fpx, fpy, fpz)
        import ctypes as C
        ox, oy, oz = np.zeros(n), np.zeros(n), np.zeros(n)
        e = [C.c_double() for _ in range(4)]
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        O.lib().orc_fnbgrp(O._p(O._i(s.ss_index)), O._p(f(s.xp0)), O._p(f(s.yp0)), O._p(f(s.zp0)),
                           O._p(f(s.charge)), O._p(typ), O._p(f(s.ecc12)), O._p(f(s.ecc6)), C.c_int32(int(ewald)),
                           C.c_double(s.alpha), O._p(O._i(ingrp)), C.c_int(ingrp.shape[1]), O._p(ox), O._p(oy),
                           O._p(oz), *[C.byref(v) for v in e])
        assert relrms(np.stack([fpx, fpy, fpz]), np.stack([ox, oy, oz])) < 1e-12
        assert rel(g[2], e[2].value) < 1e-12 and rel(g[0], e[0].value) < 1e-12


def test_decide_force():
    from orac_hip.api import Context
    for r in (0.0, 3.0, 5.1, 7.8, 10.0):
        for a in (0.3, 0.37, 0.4):
            assert Context.decide_force(r, a) == O.decide_force(r, a)


@pytest.mark.parametrize("shell", [0, 1])
def test_force_owner_form(gpu_ctx, small_solute, shell, monkeypatch):
    """ORAC_HIP_PAIR_KERNEL=own: the force entry point takes the owner-atom
    loop form (full list) for groups of at most three atoms; against the
    oracle, including force.f's massinfty rule."""
    import dataclasses
    monkeypatch.setenv("ORAC_HIP_PAIR_KERNEL", "own")
    s = small_solute
    mass = s.mass.copy()
    mass[:20] = 1e21
    s2 = dataclasses.replace(s, mass=mass)
    nnlpp, fa, _ = _oracle_list(s2)
    sh = dict(SHELLS[shell])
    if sh["rout"] is None:
        sh["rout"] = s.rcut
    g = _force_gpu(gpu_ctx, s2, nnlpp, fa, massinfty=1e20, **sh)
    o = O.force(s2, nnlpp, fa=fa, massinfty=1e20, **sh)
    assert relrms(g["f"], o["f"]) < F_TOL
    for k in ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"):
        if o[k] != 0.0:
            assert rel(g[k], o[k]) < E_TOL, k
    assert _stress_close(g["stress"], o["stress"])
