"""The oracle against golden vectors of the reference's own routines.

tests/golden/reference_vectors.npz was written by tests/golden/make_golden.py
from the reference Fortran compiled in place (oracle/_ref).  These checks pin
the oracle where that build is absent: bit-exact for lists, forces, energies,
stress, corrections, B-spline moduli and the FFTW scalar_sum; within 1e-13 for
fft_pme, whose reference build is the fftpack branch.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")
E6 = ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


@pytest.mark.parametrize("order,k", [(4, 24), (5, 20), (6, 30), (5, 27)])
def test_golden_bsp_moduli(g, order, k):
    assert np.array_equal(O.bsp_moduli(order, k), g[f"bsp_o{order}_k{k}"])


def test_golden_water_list_force(g, water216):
    s = water216
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nn = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0,
                 fa=fa)["nnlpp1"]
    assert np.array_equal(nn, g["w_nnlpp"])
    f = O.force(s, nn, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0, fa=fa)
    assert np.array_equal(f["f"], g["w_force_f"])
    assert np.array_equal(f["stress"], g["w_force_stress"])
    for k in E6:
        assert f[k] == g[f"w_force_{k}"], k
    f2 = O.force(s, nn, rinn=4.5, rout=7.0, rtolinn=0.5, rtolout=0.5, rneigh=0.3, fa=fa)
    assert np.array_equal(f2["f"], g["w_shell_f"])
    assert np.array_equal(f2["stress"], g["w_shell_stress"])
    assert np.array_equal(f2["nnlpp1"], g["w_shell_nnlpp1"])
    assert np.array_equal(np.array([f2[k] for k in E6]), g["w_shell_e"])


def test_golden_water_pme(g, water216):
    p = O.fft_pme(water216, variant="fftpack")
    assert rel(p["eer"], float(g["w_pme_eer"])) < 1e-13
    assert np.abs(p["f"] - g["w_pme_f"]).max() < 1e-13 * np.abs(g["w_pme_f"]).max()
    assert np.abs(p["stress"] - g["w_pme_stress"]).max() < 1e-13 * np.abs(g["w_pme_stress"]).max()
    assert np.array_equal(np.array(O.cself(water216)), g["w_cself"])


def test_golden_scalar_sum_fftw(g, water216):
    s = water216
    e, vir, F = O.scalar_sum(g["w_ss_fq_in"], s.alpha, s.volume, g["w_ss_recip"], [g["bsp_o5_k20"]] * 3, s.nfft)
    assert e == float(g["w_ss_eer"])
    assert np.array_equal(vir, g["w_ss_vir"])
    assert np.array_equal(F, g["w_ss_fq_out"])


def test_golden_solute_lists_force_ferrf(g, small_solute):
    t = small_solute
    rc = t.rcut + t.rtol + t.rnei
    fb = O.ForceArgs(t)
    nc = int(t.L // 3.0)
    st = O.lc_index(nc, nc, nc, rc, t.co)
    assert np.array_equal(st[0], g["s_nind"]) and np.array_equal(np.stack(st[1:]), g["s_indx"])
    nn = O.lc_list(nc, nc, nc, st, rc, t.co, fb.xpg, fb.ypg, fb.zpg)
    assert np.array_equal(nn, g["s_nnlpp"])
    f = O.force(t, nn, rinn=0.0, rout=t.rcut, rtolinn=0.0, rtolout=t.rtol, rneigh=0.0, fa=fb)
    assert np.array_equal(f["f"], g["s_force_f"])
    assert np.array_equal(f["stress"], g["s_force_stress"])
    assert np.array_equal(np.array([f[k] for k in E6]), g["s_force_e"])
    pairs = np.concatenate([t.bonds, t.bends], axis=1)
    c = O.ferrf(t, pairs, 1, 1.0)
    assert np.array_equal(np.array([c["slt"], c["slv"]]), g["s_ferrf_e"])
    assert np.array_equal(c["f"], g["s_ferrf_f"])
    c0 = O.ferrf(t, pairs, 0, 0.5)
    assert np.array_equal(np.array([c0["slt"], c0["slv"]]), g["s_ferrf_iz0_e"])
