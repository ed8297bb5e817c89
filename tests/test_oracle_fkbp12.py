"""The oracle on the reference's own test system (test1: solvated FKBP12 with
its ligand, 14,900 atoms, read from 1j4i.tpgprm and md.rst), against the
compiled reference Fortran's results on it, stored in tests/golden/fkbp12.npz
by tests/golden/make_fkbp12.py.

Serial oracle calls reproduce the Fortran's summation order, so the list,
force() and ferrf results must agree bit for bit (compared as SHA-256 digests
of the arrays' bytes); fft_pme's fftpack branch agrees to rounding.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from orac_hip import formats

E6 = ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss")


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


@pytest.fixture(scope="module")
def fk():
    s, ncell = formats.fkbp12()
    g = np.load(formats.FKBP12_PATH)
    return s, ncell, g


@pytest.fixture(scope="module")
def fk_list(fk):
    s, ncell, _ = fk
    fa = O.ForceArgs(s)
    rc = s.rcut + s.rtol + s.rnei
    st = O.lc_index(ncell, ncell, ncell, rc, s.co)
    return fa, O.lc_list(ncell, ncell, ncell, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)


def test_fixture_is_the_test1_system(fk):
    s, ncell, _ = fk
    assert s.natoms == 14900 and ncell == 18 and tuple(s.nfft) == (48, 48, 48) and s.order == 4
    # md.in: alpha 0.37, outer cutoff 10 A, update skin 1.8 A
    assert (s.alpha, s.rcut, s.rnei) == (0.37, 10.0, 1.8)
    # net charge of the neutralised system (charges are e / sqrt(unitc))
    assert abs(s.charge.sum()) < 1e-6 * np.abs(s.charge).sum()
    # rigid TIP3P waters (3 constraints each): O-H 0.9572 A, H-H 1.5139 A
    x, y, z = s.cartesian()
    c = s.constraints - 1
    d = np.sqrt((x[c[0]] - x[c[1]]) ** 2 + (y[c[0]] - y[c[1]]) ** 2 + (z[c[0]] - z[c[1]]) ** 2)
    assert c.shape[1] % 3 == 0
    assert np.all(np.minimum(np.abs(d - 0.9572), np.abs(d - 1.5139)) < 1e-3)


def test_list_matches_reference(fk, fk_list):
    _, _, g = fk
    _, nn = fk_list
    assert len(nn) == int(g["ref_nnlpp_len"])
    assert np.array_equal(digest(nn.astype(np.int32)), g["ref_nnlpp_sha"])


def test_force_matches_reference(fk, fk_list):
    s, _, g = fk
    fa, nn = fk_list
    f = O.force(s, nn, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0, fa=fa)
    assert np.array_equal(np.array([f[k] for k in E6]), g["ref_force_e"])
    assert np.array_equal(f["stress"], g["ref_force_stress"])
    assert np.array_equal(digest(f["f"]), g["ref_force_f_sha"])


def test_fft_pme_matches_reference(fk):
    s, _, g = fk
    p = O.fft_pme(s, variant="fftpack")
    assert abs(p["eer"] - float(g["ref_pme_eer"])) <= 1e-13 * abs(float(g["ref_pme_eer"]))
    assert np.abs(p["stress"] - g["ref_pme_stress"]).max() <= 1e-12 * np.abs(g["ref_pme_stress"]).max()
    fs = p["f"][:, g["ref_pme_f_idx"]]
    assert np.abs(fs - g["ref_pme_f_sample"]).max() <= 1e-11 * np.abs(g["ref_pme_f_sample"]).max()


def test_ferrf_matches_reference(fk):
    s, _, g = fk
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    c = O.ferrf(s, pairs, 1, 1.0)
    assert np.array_equal(np.array([c["slt"], c["slv"]]), g["ref_ferrf_e"])
    assert np.array_equal(digest(c["f"]), g["ref_ferrf_f_sha"])


FNB_E = ("uconf_slt", "uconf_slv", "ucoul_slt", "ucoul_slv")


@pytest.mark.parametrize("name", ["intra", "bb"])
@pytest.mark.parametrize("ewald", [1, 0])
def test_fnbgrp_matches_reference(fk, name, ewald):
    """fnbgrp on the solute's non-excluded intra-group pairs (the 1-4 pairs
    that share a group) and on the bond + 1-3 lists, with and without Ewald."""
    s, _, g = fk
    lst = g["fnb_intra"].astype(np.int32) if name == "intra" else np.concatenate([s.bonds, s.bends], axis=1)
    r = O.fnbgrp(s, lst, s.nbtype, bool(ewald))
    assert np.array_equal(np.array([r[k] for k in FNB_E]), g[f"ref_fnbgrp_{name}{ewald}_e"])
    assert np.array_equal(digest(r["f"]), g[f"ref_fnbgrp_{name}{ewald}_f_sha"])


@pytest.mark.parametrize("ewald", [1, 0])
def test_fnb14_matches_reference(fk, ewald):
    """fnb14 on the topology's 4,500 1-4 pairs (QQ-FUDGE 0.83333, LJ-FUDGE 0.5)."""
    s, _, g = fk
    r = O.fnb14(s, g["int14"].astype(np.int32), g["type14"], 0.83333, 0.5, ewald=bool(ewald), cut=True)
    assert np.array_equal(np.array([r[k] for k in FNB_E]), g[f"ref_fnb14_{ewald}_e"])
    assert np.array_equal(digest(r["f"]), g[f"ref_fnb14_{ewald}_f_sha"])


def test_fptors_matches_reference(fk):
    """fptors on the topology's 5,732 proper torsions (multiplicities 1-6, both signs)."""
    s, _, g = fk
    pt = g["potto"]
    r = O.fptors(s, g["ltor"].astype(np.int32), pt[:, 1], pt[:, 0])
    assert r["status"] == 0
    assert np.array_equal(np.array([r["utors_slt"], r["utors_slv"]]), g["ref_fptors_e"])
    assert np.array_equal(digest(r["f"]), g["ref_fptors_f_sha"])


def _erf(g):
    nbin, rlew, ruew, _ = g["erf_params"]
    return g["ref_erf_table"], (ruew - rlew) / int(nbin), rlew


@pytest.mark.parametrize("iz", [1, 0])
def test_ferrf_erf_corr_matches_reference(fk, iz):
    """ferrf's ERF_CORR branches on the reference's own correction table."""
    s, _, g = fk
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    c = O.ferrf(s, pairs, iz, 0.8, erf=_erf(g))
    assert np.array_equal(np.array([c["slt"], c["slv"]]), g[f"ref_ferrf_erf{iz}_e"])
    assert np.array_equal(digest(c["f"]), g[f"ref_ferrf_erf{iz}_f_sha"])


def test_erf_corr_table_builders_match_reference(fk):
    """The library's host restatement of int_corr_erf_spline (with de Boor's
    CUBSPL) and erf_corr_cutoff reproduces the reference's table and cutoff."""
    from orac_hip import api
    if not os.path.exists(api.library_path()):
        pytest.skip("liborac_hip.so not built")
    s, _, g = fk
    nbin, rlew, ruew, rk = g["erf_params"]
    assert api.Context.erf_corr_cutoff(s.oc, *s.nfft) == rk
    tab, tau = api.Context.int_corr_erf_spline(rlew, ruew, int(nbin), s.alpha, rk)
    assert np.array_equal(tab, g["ref_erf_table"]) and np.array_equal(tau, g["ref_erf_tau"])


def test_erf_corr_table_is_the_cutoff_integral(fk):
    """Node values are (2/pi) int_kcut^{20 alpha} exp(-k^2/4a^2) sin(kr)/(kr) dk
    (to the builder's Simpson tolerance, 1e-6) and the spline interpolates them
    with continuous first and second derivatives at the interior nodes."""
    from scipy.integrate import quad
    s, _, g = fk
    nbin, rlew, ruew, rk = g["erf_params"]
    tab, tau = g["ref_erf_table"], g["ref_erf_tau"]
    a = s.alpha
    for i in (0, 37, 150, 300):
        r = tau[i]
        v = quad(lambda k: np.exp(-k * k / (4 * a * a)) * np.sin(k * r) / (k * r), rk, 20 * a, epsabs=1e-14,
                 limit=200)[0] * 2 / np.pi
        assert abs(tab[0, i] - v) <= 2e-6 * abs(v), i
    h = np.diff(tau)[:-1]
    c = tab[:, :-2]
    val = c[0] + h * (c[1] + h * (c[2] / 2 + h * c[3] / 6))
    der = c[1] + h * (c[2] + h * c[3] / 2)
    sec = c[2] + h * c[3]
    assert np.abs(val - tab[0, 1:-1]).max() <= 1e-12 * np.abs(tab[0]).max()
    assert np.abs(der - tab[1, 1:-1]).max() <= 1e-9 * np.abs(tab[1]).max()
    assert np.abs(sec - tab[2, 1:-1]).max() <= 1e-7 * np.abs(tab[2]).max()


def test_fpbond_matches_reference(fk):
    """fpbond on the topology's 14,919 bonds (solute bonds and the waters' rigid ones)."""
    s, _, g = fk
    bp = g["bond_pots"]
    r = O.fpbond(s, g["lbnd"].astype(np.int32), bp[0], bp[1])
    assert np.array_equal(np.array([r["ubond_slt"], r["ubond_slv"]]), g["ref_fpbond_e"])
    assert np.array_equal(digest(r["f"]), g["ref_fpbond_f_sha"])


@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fpbend_matches_reference(fk, which):
    """fpbend on the topology's 3,141 bends, and on 600 of them with every
    branch set (Urey-Bradley, both coupling forms, a rest angle of pi)."""
    s, _, g = fk
    if which == "topology":
        lb, pots, key = g["lbndg"].astype(np.int32), g["bend_pots"], "ref_fpbend"
    else:
        lb, pots, key = g["lbndg"].astype(np.int32)[:, :600], g["bend_branch_pots"], "ref_fpbend_branch"
    r = O.fpbend(s, lb, list(pots))
    assert np.array_equal(np.array([r["ubend_slt"], r["ubend_slv"]]), g[key + "_e"])
    assert np.array_equal(digest(r["f"]), g[key + "_f_sha"])


@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fitors_matches_reference(fk, which):
    """fitors on the topology's 339 improper torsions (cosine, multiplicity 2)
    and on 800 quadruples with the harmonic form (rest dihedrals across
    (-pi, pi), 0 and 3.14159) and every cosine multiplicity, both signs."""
    s, _, g = fk
    if which == "topology":
        lt, pots, key = g["litr"].astype(np.int32), g["itor_pots"], "ref_fitors"
    else:
        lt, pots, key = g["ltor"].astype(np.int32)[:, :800], g["itor_branch_pots"], "ref_fitors_branch"
    r = O.fitors(s, lt, *pots)
    assert r["status"] == 0
    assert np.array_equal(np.array([r["utors_slt"], r["utors_slv"]]), g[key + "_e"])
    assert np.array_equal(digest(r["f"]), g[key + "_f_sha"])
