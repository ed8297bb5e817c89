"""The oracle's alchemical intramolecular terms (ferrf_alchemy_add,
fnb14_alchemy, fnbgrp_alchemy) and decide_force, pinned:

* bit-exact against the compiled reference where oracle/_ref is built, and
  against its results in tests/golden/alchemy_intra.npz (written by
  tests/golden/make_alchemy_intra.py) everywhere;
* against what the mathematics fixes: decide_force against an independent
  erfc (math.erfc; A&S 7.1.26 with ORAC's coefficients is within 2e-7 of it), ferrf_alchemy_add on
  two atoms in closed form and dwrk as U(lambda) - U(lambda0), fnb14_alchemy
  with no alchemical atom equal to fnb14, and its dwrk as a difference of
  two single-lambda energies.
"""
import hashlib
import math
import os
import sys

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_alchemy_intra as G  # noqa: E402
from orac_hip import synth  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden", "alchemy_intra.npz")
NAMES = [st[0] for st in G.STATES]


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


@pytest.fixture(scope="module")
def ctx():
    s, ss = G.system()
    g = np.load(GOLD)
    assert np.array_equal(digest(np.concatenate([s.xp0, s.yp0, s.zp0, s.charge])), g["inputs_sha"])
    lq = G.qq_list(s)
    p14, t14 = G.pairs14(s)
    assert np.array_equal(lq, g["listqq"]) and np.array_equal(p14, g["int14"]) and np.array_equal(t14, g["type14"])
    return s, ss, lq, p14, t14, g


def _erf(s):
    if O.ref_available():
        return G.erf_table(s)
    return None


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("intra", [0, 1])
def test_ferrf_alchemy_add_matches_reference(ctx, name, intra):
    s, ss, lq, _, _, g = ctx
    st = dict((x[0], x) for x in G.STATES)[name]
    _, lamq, _, lamq0 = G.lams(ss, st)
    r = O.ferrf_alchemy_add(s, lamq, lamq0, lq, intramol=bool(intra))
    assert np.array_equal(np.array([r["slt"], r["slv"], r["dwrk"]]), g[f"add_{name}_{intra}0_e"])
    assert np.array_equal(digest(r["f"]), g[f"add_{name}_{intra}0_f_sha"])
    erf = _erf(s)
    if erf is None:
        pytest.skip("ERF_CORR table comes from the reference build (oracle/_ref)")
    r = O.ferrf_alchemy_add(s, lamq, lamq0, lq, intramol=bool(intra), erf=erf)
    assert np.array_equal(np.array([r["slt"], r["slv"], r["dwrk"]]), g[f"add_{name}_{intra}1_e"])
    assert np.array_equal(digest(r["f"]), g[f"add_{name}_{intra}1_f_sha"])
    o = O.ref_ferrf_alchemy_add(s, lamq, lamq0, lq, intramol=bool(intra), erf=erf)
    assert np.array_equal(o["f"], r["f"]) and o["dwrk"] == r["dwrk"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("ewald", [1, 0])
def test_fnb14_fnbgrp_alchemy_match_reference(ctx, name, ewald):
    s, ss, _, p14, t14, g = ctx
    st = dict((x[0], x) for x in G.STATES)[name]
    lams = G.lams(ss, st)
    r = O.fnb14_alchemy(s, p14, t14, *G.FNB_FUDGE, lams, ewald=bool(ewald))
    assert np.array_equal(np.array([r[k] for k in O.FNB_E] + [r["dwrk"]]), g[f"fnb14_{name}_{ewald}_e"])
    assert np.array_equal(digest(r["f"]), g[f"fnb14_{name}_{ewald}_f_sha"])
    q = O.fnbgrp_alchemy(s, p14, s.nbtype, bool(ewald), lams[1])
    assert np.array_equal(np.array([q[k] for k in O.FNB_E]), g[f"fnbgrp_{name}_{ewald}_e"])
    assert np.array_equal(digest(q["f"]), g[f"fnbgrp_{name}_{ewald}_f_sha"])
    if O.ref_available():
        o = O.ref_fnb14_alchemy(s, p14, t14, *G.FNB_FUDGE, lams, ewald=bool(ewald))
        assert np.array_equal(o["f"], r["f"]) and o["dwrk"] == r["dwrk"]


def test_fnb14_alchemy_without_alchemical_atoms_is_fnb14(ctx):
    """Every lambda zero: both branches are fnb14.f's arithmetic (factors 1,
    no erf term).  Energies and the plain branch's forces agree bit for bit;
    the Ewald branch groups furpar*((erfc + ...)*rspqi) where fnb14.f groups
    (furpar*(erfc + ...))*rspqi (fnb14_alchemy.f:313-315), so its forces
    agree to rounding.  dwrk is 0."""
    s, _, _, p14, t14, _ = ctx
    z = np.zeros(s.natoms)
    for ew in (True, False):
        r = O.fnb14_alchemy(s, p14, t14, *G.FNB_FUDGE, (z, z, z, z), ewald=ew)
        p = O.fnb14(s, p14, t14, *G.FNB_FUDGE, ewald=ew)
        if ew:
            assert np.abs(r["f"] - p["f"]).max() <= 4e-16 * np.abs(p["f"]).max()
        else:
            assert np.array_equal(r["f"], p["f"])
        assert r["dwrk"] == 0.0
        assert all(r[k] == p[k] for k in O.FNB_E)


def test_fnb14_alchemy_dwrk_is_an_energy_difference(ctx):
    """Ewald branch: dwrk = U(lambda) - U(lambda0), where U(lambda0) is a call
    with lambda0 in place of lambda (same zero/equality pattern)."""
    s, ss, _, p14, t14, _ = ctx
    st = dict((x[0], x) for x in G.STATES)["mid"]
    lam, lamq, lam0, lamq0 = G.lams(ss, st)
    r = O.fnb14_alchemy(s, p14, t14, *G.FNB_FUDGE, (lam, lamq, lam0, lamq0))
    r0 = O.fnb14_alchemy(s, p14, t14, *G.FNB_FUDGE, (lam0, lamq0, lam0, lamq0))
    u = sum(r[k] for k in O.FNB_E)
    u0 = sum(r0[k] for k in O.FNB_E)
    assert abs(r["dwrk"] - (u - u0)) <= 1e-12 * (abs(u) + abs(u0))


def test_ferrf_alchemy_add_two_atoms_closed_form():
    """One added-removed pair (facQ = 0): E = -(1 - |lq_a|)(1 - |lq_b|) q_a q_b
    erf(alpha r)/r with the A&S erf; the force is along the pair and equal and
    opposite; dwrk = E(lambdaq) - E(lambdaq0)."""
    import dataclasses
    s, _ = G.system()
    xyz = np.zeros((3, s.natoms))
    xyz[:, 1] = [1.3, -0.7, 2.1]
    co = np.asarray(s.co).reshape(3, 3, order="F")
    red = np.linalg.solve(co, xyz)
    t = dataclasses.replace(s, xp0=red[0].copy(), yp0=red[1].copy(), zp0=red[2].copy())
    lq = np.zeros(s.natoms)
    lq0 = np.zeros(s.natoms)
    lq[0], lq[1], lq0[0], lq0[1] = 0.3, -0.6, 0.35, -0.5
    r = O.ferrf_alchemy_add(t, lq, lq0, np.array([[1], [2], [0]], dtype=np.int32))
    x, y, z = t.cartesian()
    rr = math.sqrt((x[0] - x[1]) ** 2 + (y[0] - y[1]) ** 2 + (z[0] - z[1]) ** 2)
    qa, qb = s.charge[0], s.charge[1]

    def energy(la, lb):
        return -(1 - abs(la)) * (1 - abs(lb)) * qa * qb * (1 - O_erfc_as(s.alpha * rr)) / rr
    e = energy(0.3, -0.6)
    assert abs(r["slt"] - e) <= 1e-13 * abs(e) and r["slv"] == 0.0
    assert abs(r["dwrk"] - (e - energy(0.35, -0.5))) <= 1e-13 * abs(e)
    f = r["f"]
    assert np.allclose(f[:, 0], -f[:, 1], rtol=0, atol=1e-15 * np.abs(f).max())
    sep = np.array([x[0] - x[1], y[0] - y[1], z[0] - z[1]])
    assert np.linalg.norm(np.cross(f[:, 0], sep)) <= 1e-12 * np.linalg.norm(f[:, 0]) * rr
    assert np.abs(f[:, 2:]).max() == 0.0


def O_erfc_as(x):
    """A&S 7.1.26 as an independent Python expression (for the closed forms)."""
    a1, a2, a3, a4, a5, p = 0.2548296, -0.28449674, 1.4214137, -1.453152, 1.0614054, 0.3275911
    t = 1.0 / (1.0 + p * x)
    return ((((a5 * t + a4) * t + a3) * t + a2) * t + a1) * t * math.exp(-x * x)


def test_ferrf_alchemy_add_intramol_keeps_only_mixed_pairs(ctx):
    """With intramol the facQ = 1 pairs are skipped: the result equals a call
    on the facQ = 0 pairs alone."""
    s, ss, lq, _, _, _ = ctx
    _, lamq, _, lamq0 = G.lams(ss, G.STATES[0])
    a = O.ferrf_alchemy_add(s, lamq, lamq0, lq, intramol=True)
    b = O.ferrf_alchemy_add(s, lamq, lamq0, np.ascontiguousarray(lq[:, lq[2] == 0]), intramol=False)
    assert np.array_equal(a["f"], b["f"]) and a["dwrk"] == b["dwrk"]


def test_decide_force_matches_reference_at_the_threshold(ctx):
    """Points within a few ulps of erfc_AS(alpha r) = 1e-3, where the
    association of alpha*r and of the exponent's argument decides the result."""
    g = ctx[5]
    got = np.array([O.decide_force(r, a) for r, a in g["decide_pts"]], dtype=np.int8)
    assert np.array_equal(got, g["decide_ref"])
    assert 0 < g["decide_ref"].sum() < len(g["decide_ref"])
    if O.ref_available():
        assert all(O.ref_decide_force(r, a) == bool(v) for (r, a), v in zip(g["decide_pts"], got))


@pytest.mark.parametrize("alpha", G.ALPHAS)
def test_decide_force_against_math_erfc(alpha):
    """lskip_ewald is true exactly where erfc(alpha r) < 1e-3, up to the A&S
    error (|erfc_AS - erfc| <= 2e-7 with the coefficients as ORAC writes
    them; 1.5e-7 is the textbook bound for the full-precision ones): a wrong
    threshold (1e-4), a dropped exponent or a missing factor alpha moves the
    boundary far outside that."""
    for x in np.linspace(0.0, 4.0, 4001):
        e = math.erfc(x)
        r = x / alpha
        if e < 1e-3 - 3e-7:
            assert O.decide_force(r, alpha), (x, e)
        elif e > 1e-3 + 3e-7:
            assert not O.decide_force(r, alpha), (x, e)
    for x in np.linspace(0.0, 4.0, 401):
        assert abs(O_erfc_as(x) - math.erfc(x)) <= 2e-7


def test_qq_list_builder():
    """initialize_lambda.f:66-140: added atoms first, then removed; bonded,
    1-3 and 1-4 pairs skipped; facQ 1 within a kind, 0 across."""
    lst = synth.alchemy_qq_list([3, 1], [5], (np.array([[1], [3]]),))
    assert lst.tolist() == [[3, 1], [5, 5], [0, 0]]
    lst = synth.alchemy_qq_list([1, 2], [4, 6], ())
    assert lst.shape == (3, 6)
    assert lst[2].tolist() == [1, 0, 0, 0, 0, 1]
