"""The C oracle against the reference's own Fortran, compiled in place (oracle/_ref).

Skipped where the reference library was not built (no reference tree).  The
serial oracle reproduces the serial reference bit for bit for every routine
whose arithmetic is deterministic; the PME chain is compared on the
reference's fftpack branch (the FFTW branch needs FFTW, absent here), whose
FFT differs from ours only by rounding.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel, relrms

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")

SHELLS = [
    dict(rinn=0.0, rout=None, rtolinn=0.0, rtolout=0.3, rneigh=0.0),        # single m-shell
    dict(rinn=4.0, rout=7.0, rtolinn=0.5, rtolout=0.5, rneigh=0.6),        # switched inner and outer
    dict(rinn=2.5, rout=5.1, rtolinn=0.3, rtolout=0.3, rneigh=0.4),
]


def _lists(s):
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    if s.L > 3 * rc:
        nc = int(s.L // 3.0)
        st = O.lc_index(nc, nc, nc, rc, s.co)
        return O.lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg), fa
    u = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0, fa=fa)
    return u["nnlpp1"], fa


@pytest.mark.parametrize("sysname", ["water216", "small_solute"])
def test_bsp_moduli_exact(sysname, request):
    s = request.getfixturevalue(sysname)
    ref = O.ref_bsp_moduli(s.order, s.nfft)
    for k, r in zip(s.nfft, ref):
        assert np.array_equal(O.bsp_moduli(s.order, k), r)


def test_bsp_moduli_odd_even_orders():
    for order in (4, 5, 6, 7):
        for k in (16, 20, 24, 30):
            ref = O.ref_bsp_moduli(order, (k, k, k))[0]
            assert np.array_equal(O.bsp_moduli(order, k), ref), (order, k)


@pytest.mark.parametrize("sysname", ["water216", "small_solute"])
def test_fft_pme_fftpack_branch(sysname, request):
    s = request.getfixturevalue(sysname)
    o = O.fft_pme(s, variant="fftpack")
    r = O.ref_fft_pme(s)
    assert rel(o["eer"], r["eer"]) < 1e-13
    assert relrms(o["f"], r["f"]) < 1e-13
    assert np.abs(o["stress"] - r["stress"]).max() < 1e-13 * np.abs(r["stress"]).max()


@pytest.mark.parametrize("sysname", ["water216", "small_solute"])
def test_scalar_sum_fftw_variant_exact(sysname, request):
    """The half-complex scalar_sum with its virial masks, on the same FQ."""
    s = request.getfixturevalue(sysname)
    x, y, z = s.cartesian()
    recip = O.recip_of(s.oc)
    Q = O.charge_grid(x, y, z, s.charge, recip, s.nfft, s.order)["Q"]
    FQ = np.fft.rfftn(Q)
    bsp = [O.bsp_moduli(s.order, k) for k in s.nfft]
    e1, v1, F1 = O.scalar_sum(FQ, s.alpha, s.volume, recip, bsp, s.nfft)
    e2, v2, F2 = O.ref_scalar_sum_fftw(FQ, s.alpha, s.volume, recip, bsp, s.nfft)
    assert e1 == e2
    assert np.array_equal(v1, v2)
    assert np.array_equal(F1, F2)


def test_fftw_and_fftpack_branches_differ_by_real4_pi(water216):
    """The FFTW half-complex sum weights each mirror pair by 2 with the +m
    modulus; DFTMOD's REAL(4) 2*pi makes |b(m)|^2 and |b(-m)|^2 differ at
    ~1e-7, so the two branches of the reference disagree at ~1e-8."""
    a = O.fft_pme(water216, variant="fftw")
    b = O.fft_pme(water216, variant="fftpack")
    assert 1e-10 < rel(a["eer"], b["eer"]) < 1e-6


@pytest.mark.parametrize("sysname", ["water216", "small_solute"])
def test_lists_exact(sysname, request):
    s = request.getfixturevalue(sysname)
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    if s.L > 3 * rc:
        nc = int(s.L // 3.0)
        so, sr = O.lc_index(nc, nc, nc, rc, s.co), O.ref_lc_index(nc, nc, nc, rc, s.co)
        assert all(np.array_equal(a, b) for a, b in zip(so, sr))
        lo = O.lc_list(nc, nc, nc, so, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
        lr = O.ref_lc_list(nc, nc, nc, sr, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
    else:
        allg = O.all_groups_list(s.ngrp)
        lo = O.force(s, allg, rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0, fa=fa)["nnlpp1"]
        lr = O.ref_force(s, allg, rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0, fa=fa)["nnlpp1"]
    assert np.array_equal(lo, lr)


@pytest.mark.parametrize("shell", range(len(SHELLS)))
@pytest.mark.parametrize("sysname", ["water216", "small_solute"])
def test_force_exact(sysname, shell, request):
    s = request.getfixturevalue(sysname)
    nnlpp, fa = _lists(s)
    sh = dict(SHELLS[shell])
    if sh["rout"] is None:
        sh["rout"] = s.rcut
    o = O.force(s, nnlpp, fa=fa, **sh)
    r = O.ref_force(s, nnlpp, fa=fa, **sh)
    for k in ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss", "ncount1", "rtolinn"):
        assert o[k] == r[k], k
    assert np.array_equal(o["f"], r["f"])
    assert np.array_equal(o["stress"], r["stress"])
    assert np.array_equal(o["nnlpp1"], r["nnlpp1"])


def test_force_lskip_and_plain_coulomb(small_solute):
    s = small_solute
    nnlpp, fa = _lists(s)
    for lskip, ewald in ((True, True), (True, False)):
        o = O.force(s, nnlpp, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0,
                    lskip_ewald=lskip, ewald=ewald)
        r = O.ref_force(s, nnlpp, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0,
                        lskip_ewald=lskip, ewald=ewald)
        assert np.array_equal(o["f"], r["f"])
        assert o["ucoul_ss"] == r["ucoul_ss"] and o["uconf_slt"] == r["uconf_slt"]


def test_force_openmp_variant_close(small_solute):
    s = small_solute
    nnlpp, fa = _lists(s)
    r = O.ref_force(s, nnlpp, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0)
    o = O.force(s, nnlpp, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0, nthreads=4)
    assert relrms(o["f"], r["f"]) < 1e-14
    assert rel(o["ucoul_slv"], r["ucoul_slv"]) < 1e-13


@pytest.mark.parametrize("which", ["constraints", "bonds", "bends"])
def test_ferrf_exact(small_solute, which):
    s = small_solute
    pairs = getattr(s, which)
    for iz, fudge in ((1, 1.0), (1, 0.8333), (0, 1.0)):
        a = O.ferrf(s, pairs, iz, fudge)
        b = O.ref_ferrf(s, pairs, iz, fudge)
        assert a["slt"] == b["slt"] and a["slv"] == b["slv"]
        assert np.array_equal(a["f"], b["f"])


def test_cself_exact(small_solute):
    for kcut in (1e9, 3.0):
        assert O.cself(small_solute, kcut) == O.ref_cself(small_solute, kcut)


@pytest.mark.skipif(not O.ref_omp_available(), reason="oracle/_ref/libref_omp.so not built (no reference tree)")
def test_reference_openmp_build_matches_serial():
    """The reference's OpenMP build (ORAC's per-thread lists and contact
    maps, timed as the CPU baseline) gives the serial build's forces on the
    24,460-atom system, to the rounding of its thread reductions."""
    from orac_hip import synth
    s = synth.config("solvated24k")
    corr = np.concatenate([s.bonds, s.bends], axis=1)
    fa = O.ForceArgs(s)
    rc = s.rcut + s.rtol + s.rnei
    nc = int(s.L // 3.0)
    st = O.ref_lc_index(nc, nc, nc, rc, s.co)
    nn = O.ref_lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
    d = O.ref_force(s, nn, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0, fa=fa)
    fr = d["f"] + O.ref_fft_pme(s)["f"] + O.ref_ferrf(s, corr, 1, 1.0)["f"]
    for nthr in (1, 3):
        f, done, _ = O.ref_omp_evaluation(s, corr, nthr, fa=fa)
        assert done == 1
        assert np.sqrt(((f - fr) ** 2).sum() / (fr ** 2).sum()) < 1e-13
