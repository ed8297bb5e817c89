"""The GPU path on the reference's own test system (test1: solvated FKBP12 with
its ligand, 14,900 atoms) against the compiled reference Fortran's results
(tests/golden/fkbp12.npz) and the oracle.

Plain entry points: the list must be the reference's bit for bit (SHA-256 of
its bytes); force() and fft_pme within the FP64 tolerances (forces relative
RMS <= 1e-9, energies and virial <= 1e-10).  The resident path (list, pairs,
PME and corrections in one evaluation) is checked in FP64 and in mixed
precision (forces relative RMS <= 5e-6).
"""
import hashlib

import numpy as np
import pytest

import oracle as O
from conftest import rel, relrms

pytestmark = pytest.mark.gpu

E6 = ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss")


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


@pytest.fixture(scope="module")
def fk():
    from orac_hip import formats
    s, ncell = formats.fkbp12()
    return s, ncell, np.load(formats.FKBP12_PATH)


@pytest.fixture(scope="module")
def fk_list(fk):
    s, ncell, _ = fk
    fa = O.ForceArgs(s)
    rc = s.rcut + s.rtol + s.rnei
    st = O.lc_index(ncell, ncell, ncell, rc, s.co)
    return fa, st, O.lc_list(ncell, ncell, ncell, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)


def test_lc_list_is_reference(gpu_ctx, fk, fk_list):
    s, nc, g = fk
    fa, _, _ = fk_list
    rc = s.rcut + s.rtol + s.rnei
    nind, ii, jj, kk = gpu_ctx.lc_index(2500, nc, nc, nc, rc, s.co)
    nn = np.zeros(int(g["ref_nnlpp_len"]) + 1024, np.int32)
    n = gpu_ctx.lc_list(nc, nc, nc, nind, ii, jj, kk, rc, s.co, fa.xpg, fa.ypg, fa.zpg, s.ngrp, nn)
    assert n == int(g["ref_nnlpp_len"])
    assert np.array_equal(digest(nn[:n]), g["ref_nnlpp_sha"])


def test_force_matches_reference(gpu_ctx, fk, fk_list):
    from test_gpu_parity import _force_gpu, _stress_close
    s, _, g = fk
    fa, _, nn = fk_list
    kw = dict(rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0)
    r = _force_gpu(gpu_ctx, s, nn, fa, **kw)
    o = O.force(s, nn, fa=fa, **kw)
    assert np.array_equal(digest(o["f"]), g["ref_force_f_sha"])  # the oracle is the reference here
    err = relrms(r["f"], o["f"])
    print("fkbp12 force relrms", err)
    assert err < 1e-9
    for k, e in zip(E6, g["ref_force_e"]):
        assert rel(r[k], float(e)) < 1e-10, k
    assert _stress_close(r["stress"], g["ref_force_stress"])


def test_fft_pme_matches_reference(gpu_ctx, fk):
    """The GPU follows the reference's FFTW branch (the oracle's 'fftw'
    variant) to the FP64 tolerances.  The stored reference values come from
    its fftpack branch, which differs from the FFTW branch by ~3e-8 in energy
    and ~1e-7 in stress on this system (DFTMOD's REAL(4) 2 pi makes |b(m)|^2
    and |b(-m)|^2 differ; the FFTW branch weights mirror pairs by the +m
    modulus), so against those the bound is 3e-7."""
    from test_gpu_parity import _pme_gpu, _stress_close
    s, _, g = fk
    r = _pme_gpu(gpu_ctx, s)
    o = O.fft_pme(s, nthreads=4)
    print("fkbp12 eer", r["eer"], o["eer"], float(g["ref_pme_eer"]))
    assert rel(r["eer"], o["eer"]) < 1e-10
    assert _stress_close(r["stress"], o["stress"])
    assert relrms(r["f"], o["f"]) < 1e-9
    assert rel(r["eer"], float(g["ref_pme_eer"])) < 3e-7
    assert _stress_close(r["stress"], g["ref_pme_stress"], tol=3e-7)
    assert relrms(r["f"][:, g["ref_pme_f_idx"]], g["ref_pme_f_sample"]) < 3e-7


def test_ferrf_matches_reference(gpu_ctx, fk):
    s, _, g = fk
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.ferrf(s.ss_index, s.alpha, s.charge, 1.0, x, y, z, 1, pairs, pairs.shape[1], fpx, fpy, fpz)
    assert rel(a, float(g["ref_ferrf_e"][0])) < 1e-10
    assert b == float(g["ref_ferrf_e"][1]) == 0.0  # the solute's bonds and 1-3 pairs are solute-solute
    o = O.ferrf(s, pairs, 1, 1.0)
    assert np.array_equal(digest(o["f"]), g["ref_ferrf_f_sha"])
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-9


@pytest.mark.parametrize("precision", [0, 1])
def test_resident_matches_oracle(gpu_ctx, fk, precision):
    from orac_hip.api import ResidentSystem
    from test_gpu_resident import oracle_total
    s, nc, _ = fk
    cp = np.concatenate([s.bonds, s.bends], axis=1)
    cf = np.ones(cp.shape[1])
    rs = ResidentSystem(gpu_ctx, s, precision=precision, ncell=nc, corr_pairs=cp, corr_fudge=cf)
    rs.eval(rebuild_list=True)
    r = rs.results()
    o = oracle_total(s, cp, cf)
    err = relrms(r["f"], o["f"])
    print("fkbp12 resident precision", precision, "f relrms", err, "eer", rel(r["eer"], o["eer"]))
    if precision == 0:
        assert err < 1e-9
        assert rel(r["eer"], o["eer"]) < 1e-10
        for k in E6:
            if o["d"][k] != 0.0:
                assert rel(r[k], o["d"][k]) < 1e-10, k
        assert rel(r["corr_slt"], o["corr_slt"]) < 1e-10
        assert np.abs(r["stress"] - o["stress"]).max() < 1e-10 * np.abs(o["stress"]).max()
    else:
        assert err < 5e-6
        assert rel(r["eer"], o["eer"]) < 1e-5
    rs.close()


@pytest.mark.parametrize("name", ["intra", "bb"])
def test_fnbgrp_matches_oracle(gpu_ctx, fk, name):
    s, _, g = fk
    lst = g["fnb_intra"].astype(np.int32) if name == "intra" else np.concatenate([s.bonds, s.bends], axis=1)
    lst = np.asfortranarray(lst)
    x, y, z = s.cartesian()
    n = s.natoms
    for ewald in (True, False):
        fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
        e = gpu_ctx.fnbgrp(s.ss_index, x, y, z, s.charge, s.nbtype, s.ecc12, s.ecc6, ewald, s.alpha, lst,
                           lst.shape[1], fpx, fpy, fpz)
        o = O.fnbgrp(s, lst, s.nbtype, ewald)
        assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
        for got, k in zip(e, ("uconf_slt", "uconf_slv", "ucoul_slt", "ucoul_slv")):
            assert (got == 0.0 and o[k] == 0.0) or rel(got, o[k]) < 1e-12, k


@pytest.mark.parametrize("ewald", [True, False])
def test_fnb14_matches_oracle(gpu_ctx, fk, ewald):
    s, _, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    i14 = np.asfortranarray(g["int14"].astype(np.int32))
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    e = gpu_ctx.fnb14(s.ss_index, x, y, z, s.charge, n, s.ecc12, s.ecc6, True, ewald, s.alpha, i14, i14.shape[1],
                      g["type14"], 0.83333, 0.5, fpx, fpy, fpz)
    o = O.fnb14(s, i14, g["type14"], 0.83333, 0.5, ewald=ewald, cut=True)
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    for got, k in zip(e, ("uconf_slt", "uconf_slv", "ucoul_slt", "ucoul_slv")):
        assert (got == 0.0 and o[k] == 0.0) or rel(got, o[k]) < 1e-12, k


def test_fptors_matches_oracle(gpu_ctx, fk):
    s, _, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    lt = np.asfortranarray(g["ltor"].astype(np.int32))
    pt = g["potto"]
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.fptors(s.ss_index, lt, lt.shape[1], x, y, z, pt[:, 1], pt[:, 0], fpx, fpy, fpz, n)
    o = O.fptors(s, lt, pt[:, 1], pt[:, 0])
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    assert rel(a, o["utors_slt"]) < 1e-12 and b == o["utors_slv"] == 0.0


def test_fptors_every_multiplicity(gpu_ctx):
    """Synthetic torsions of every multiplicity the reference knows (1-8, 12, 14), both signs."""
    from orac_hip import synth
    s = synth.water_box()
    rng = np.random.default_rng(11)
    mults = [1, 2, 3, 4, 5, 6, 7, 8, 12, 14]
    nt = 2 * len(mults)
    lt = np.asfortranarray(np.arange(4 * nt, dtype=np.int32).reshape(nt, 4).T + 1)
    pota = np.array([sg * m for m in mults for sg in (1, -1)], float)
    potb = rng.uniform(0.2, 3.0, nt)
    x, y, z = (rng.normal(0.0, 1.5, s.natoms) for _ in range(3))
    n = s.natoms
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.fptors(s.ss_index, lt, nt, x, y, z, pota, potb, fpx, fpy, fpz, n)
    o = O.fptors(s, lt, pota, potb, coords=(x, y, z))
    assert o["status"] == 0
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    assert rel(a + b, o["utors_slt"] + o["utors_slv"]) < 1e-12


@pytest.mark.parametrize("iz", [1, 0])
def test_ferrf_erf_corr_matches_oracle(gpu_ctx, fk, iz):
    """ferrf with ERF_CORR on the reference's correction table."""
    s, _, g = fk
    nbin, rlew, ruew, _ = g["erf_params"]
    tab, delew = g["ref_erf_table"], (ruew - rlew) / int(nbin)
    pairs = np.asfortranarray(np.concatenate([s.bonds, s.bends], axis=1))
    x, y, z = s.cartesian()
    n = s.natoms
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.ferrf(s.ss_index, s.alpha, s.charge, 0.8, x, y, z, iz, pairs, pairs.shape[1], fpx, fpy, fpz,
                         erf_corr=True, erf_arr_corr=tab, delew=delew, rlew=rlew)
    o = O.ferrf(s, pairs, iz, 0.8, erf=(tab, delew, rlew))
    assert rel(a, o["slt"]) < 1e-12 and b == o["slv"]
    if iz == 1:
        assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12


def test_fpbond_matches_oracle(gpu_ctx, fk):
    s, _, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    lb = np.asfortranarray(g["lbnd"].astype(np.int32))
    bp = g["bond_pots"]
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.fpbond(s.ss_index, lb, lb.shape[1], x, y, z, bp[0], bp[1], fpx, fpy, fpz, n)
    o = O.fpbond(s, lb, bp[0], bp[1])
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    assert rel(a, o["ubond_slt"]) < 1e-12 and rel(b, o["ubond_slv"]) < 1e-9


@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fpbend_matches_oracle(gpu_ctx, fk, which):
    """fpbend on the topology's bends and on bends with every branch set."""
    s, _, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    lb = g["lbndg"].astype(np.int32)
    pots = g["bend_pots"]
    if which == "branch":
        lb, pots = lb[:, :600], g["bend_branch_pots"]
    lb = np.asfortranarray(lb)
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.fpbend(s.ss_index, lb, lb.shape[1], x, y, z, *pots, fpx, fpy, fpz, n)
    o = O.fpbend(s, lb, list(pots))
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    assert rel(a, o["ubend_slt"]) < 1e-12 and b == o["ubend_slv"] == 0.0


@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fitors_matches_oracle(gpu_ctx, fk, which):
    """fitors on the topology's impropers and on impropers of every form."""
    s, _, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    if which == "topology":
        lt, pots = g["litr"].astype(np.int32), g["itor_pots"]
    else:
        lt, pots = g["ltor"].astype(np.int32)[:, :800], g["itor_branch_pots"]
    lt = np.asfortranarray(lt)
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b = gpu_ctx.fitors(s.ss_index, lt, lt.shape[1], x, y, z, *pots, fpx, fpy, fpz, n)
    o = O.fitors(s, lt, *pots)
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    assert rel(a, o["utors_slt"]) < 1e-12 and b == o["utors_slv"] == 0.0
