"""REM segment scaling of the direct-space force (force_REM.f): the oracle's
restatement against the reference compiled in place, bit for bit, on the
seg_select kinds (rem.f90:631-645), switched inner and outer shells, the
lskip_ewald and plain-Coulomb branches (including force_REM.f:667, whose
Coulomb term reads the distance an earlier pair or group left in the
routine's rsp), and lskip_force with ljintra (charges off).  Plus what the
mathematics fixes: rem_f3 = 1 gives force.f's forces and energies exactly,
and the segment energy is the force.f energy of the selected pairs."""
import numpy as np
import pytest

import oracle as O
from conftest import rel

E6 = ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss")
SHELLS = [
    dict(rinn=0.0, rout=None, rtolinn=0.0, rtolout=0.3, rneigh=0.0),
    dict(rinn=4.0, rout=7.0, rtolinn=0.5, rtolout=0.5, rneigh=0.6),
]


def _lists(s):
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nc = int(s.L // 3.0)
    st = O.lc_index(nc, nc, nc, rc, s.co)
    return O.lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg), fa


def _segment(s, seed=3):
    """A segment: the solute's first half plus a few waters."""
    rng = np.random.default_rng(seed)
    rem = np.zeros(s.natoms, np.int32)
    nsol = int((np.asarray(s.ss_index) == 1).sum())
    rem[: nsol // 2] = 1
    rem[nsol + 3 * rng.choice((s.natoms - nsol) // 3, 5, replace=False)] = 1
    return rem


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")
@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("shell", [0, 1])
def test_force_rem_exact(small_solute, kind, shell):
    s = small_solute
    nnlpp, fa = _lists(s)
    rem = _segment(s)
    sh = dict(SHELLS[shell])
    if sh["rout"] is None:
        sh["rout"] = s.rcut
    for lskip, ewald in ((False, True), (True, True), (True, False)):
        kw = dict(rem_kind=kind, rem_f3=0.55, lskip_ewald=lskip, ewald=ewald, fa=fa, **sh)
        o = O.force_rem(s, nnlpp, rem, **kw)
        r = O.ref_force_rem(s, nnlpp, rem, **kw)
        for k in E6 + ("ncount1",):
            assert o[k] == r[k], (k, lskip, ewald)
        assert np.array_equal(o["f"], r["f"]), (lskip, ewald)
        assert np.array_equal(o["stress"], r["stress"])
        assert np.array_equal(o["nnlpp1"], r["nnlpp1"])
        if not (lskip and ewald):
            # (with lskip_ewald and Ewald on, no pair sets the reference's
            # local coul, which its segment energy adds: undefined there)
            assert o["seg_ene"] == r["seg_ene"], (lskip, ewald)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")
def test_force_rem_charges_off(small_solute):
    """lskip_force: without ljintra nothing is computed; with it the charges count as zero."""
    s = small_solute
    nnlpp, fa = _lists(s)
    rem = _segment(s)
    kw = dict(rem_kind=0, rem_f3=0.7, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0, fa=fa)
    o = O.force_rem(s, nnlpp, rem, lskip_force=True, ljintra=True, **kw)
    r = O.ref_force_rem(s, nnlpp, rem, lskip_force=True, ljintra=True, **kw)
    assert np.array_equal(o["f"], r["f"]) and o["seg_ene"] == r["seg_ene"]
    assert o["ucoul_slt"] == 0.0 and o["uconf_slt"] == r["uconf_slt"]
    o = O.force_rem(s, nnlpp, rem, lskip_force=True, ljintra=False, **kw)
    assert not o["f"].any() and o["seg_ene"] == 0.0


@pytest.mark.parametrize("shell", [0, 1])
def test_force_rem_unit_factor_is_force(small_solute, shell):
    """rem_f3 = 1: forces, energies and virial are force.f's exactly (every
    scaling multiplies by one); the segment energy equals force.f's energy
    over the selected pairs, summed here from kind-0 minus kind-... splits:
    the three kinds partition as  kind0 = kind1 + kind2."""
    s = small_solute
    nnlpp, fa = _lists(s)
    rem = _segment(s)
    sh = dict(SHELLS[shell])
    if sh["rout"] is None:
        sh["rout"] = s.rcut
    f = O.force(s, nnlpp, fa=fa, **sh)
    seg = {}
    for kind in (0, 1, 2):
        o = O.force_rem(s, nnlpp, rem, rem_kind=kind, rem_f3=1.0, fa=fa, **sh)
        assert np.array_equal(o["f"], f["f"]) and np.array_equal(o["stress"], f["stress"])
        for k in E6:
            assert rel(o[k], f[k]) < 1e-14, k
        seg[kind] = o["seg_ene"]
    assert abs(seg[0] - (seg[1] + seg[2])) <= 1e-12 * (abs(seg[1]) + abs(seg[2]))
    # no segment atom: nothing selected; every atom: kind 1 selects every pair
    none = O.force_rem(s, nnlpp, np.zeros(s.natoms, np.int32), rem_kind=0, rem_f3=0.3, fa=fa, **sh)
    assert none["seg_ene"] == 0.0 and np.array_equal(none["f"], f["f"])
    allin = O.force_rem(s, nnlpp, np.ones(s.natoms, np.int32), rem_kind=1, rem_f3=1.0, fa=fa, **sh)
    etot = sum(f[k] for k in E6)
    assert rel(allin["seg_ene"], etot) < 1e-12


def test_force_rem_scales_segment_forces(small_solute):
    """With every atom in the segment (kind 1) and no switching, the forces
    are force.f's times rem_f3 and the energies unscaled."""
    s = small_solute
    nnlpp, fa = _lists(s)
    sh = dict(rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.0, rneigh=0.0)
    f = O.force(s, nnlpp, fa=fa, **sh)
    o = O.force_rem(s, nnlpp, np.ones(s.natoms, np.int32), rem_kind=1, rem_f3=0.5, fa=fa, **sh)
    assert np.abs(o["f"] - 0.5 * f["f"]).max() <= 1e-13 * np.abs(f["f"]).max()
    for k in E6:
        assert rel(o[k], f[k]) < 1e-14, k


# ---------------------------------------------------------------- bonded REM twins

@pytest.fixture(scope="module")
def fk():
    from orac_hip import formats
    s, _ = formats.fkbp12()
    return s, np.load(formats.FKBP12_PATH)


def _rem_flags(n, seed):
    return (np.random.default_rng(seed).random(n) < 0.4).astype(np.int32)


def _branch_pots(pt, n):
    """Proper torsions with every multiplicity fptors_rem evaluates, both signs,
    including the 60-degree-phase forms 12 and 14 that the REM twin treats
    apart (no sign factor, dsoa = 0 at |sin| < 1e-8)."""
    mult = np.array([1, 2, 3, 4, 5, 6, 7, 8, 12, 14, -1, -2, -3, -4, -6, -12, -14], float)
    return mult[np.arange(n) % len(mult)], pt[:n, 0]


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")
@pytest.mark.parametrize("ewald", [1, 0])
def test_fnb14_rem_exact(fk, ewald):
    s, g = fk
    i14 = g["int14"].astype(np.int32)
    fl = _rem_flags(i14.shape[1], 1)
    o = O.fnb14_rem(s, i14, g["type14"], 0.83333, 0.5, fl, 0.61, ewald=bool(ewald))
    r = O.ref_fnb14_rem(s, i14, g["type14"], 0.83333, 0.5, fl, 0.61, ewald=bool(ewald))
    for k in O.FNB_E + ("seg_ene",):
        assert o[k] == r[k], k
    assert np.array_equal(o["f"], r["f"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")
@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fptors_rem_exact(fk, which):
    s, g = fk
    lt, pt = g["ltor"].astype(np.int32), g["potto"]
    pa, pb = pt[:, 1], pt[:, 0]
    if which == "branch":
        lt = lt[:, :1700]
        pa, pb = _branch_pots(pt, 1700)
    fl = _rem_flags(lt.shape[1], 2)
    o = O.fptors_rem(s, lt, pa, pb, fl, 0.47)
    r = O.ref_fptors_rem(s, lt, pa, pb, fl, 0.47)
    assert o["status"] == 0
    for k in ("utors_slt", "utors_slv", "seg_ene"):
        assert o[k] == r[k], k
    assert np.array_equal(o["f"], r["f"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")
@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fitors_rem_exact(fk, which):
    s, g = fk
    if which == "topology":
        lt, pots = g["litr"].astype(np.int32), g["itor_pots"]
    else:
        lt, pots = g["ltor"].astype(np.int32)[:, :800], g["itor_branch_pots"]
    fl = _rem_flags(lt.shape[1], 3)
    o = O.fitors_rem(s, lt, *pots, fl, 0.52)
    r = O.ref_fitors_rem(s, lt, *pots, fl, 0.52)
    assert o["status"] == 0
    for k in ("utors_slt", "utors_slv", "seg_ene"):
        assert o[k] == r[k], k
    assert np.array_equal(o["f"], r["f"])


def test_bonded_rem_unit_factor_and_segment_energy(fk):
    """f2 = 1 leaves fnb14, fitors and (outside the 12/14 forms) fptors
    exactly as the plain routines; no flag gives seg_ene = 0; every flag gives
    the total energy as seg_ene and forces scaled by f2."""
    s, g = fk
    i14, t14 = g["int14"].astype(np.int32), g["type14"]
    lt, pt = g["ltor"].astype(np.int32), g["potto"]
    li, ip = g["litr"].astype(np.int32), g["itor_pots"]
    cases = [
        (lambda fl, f2: O.fnb14_rem(s, i14, t14, 0.83333, 0.5, fl, f2),
         O.fnb14(s, i14, t14, 0.83333, 0.5), i14.shape[1], ("uconf_slt", "uconf_slv", "ucoul_slt", "ucoul_slv")),
        (lambda fl, f2: O.fptors_rem(s, lt, pt[:, 1], pt[:, 0], fl, f2),
         O.fptors(s, lt, pt[:, 1], pt[:, 0]), lt.shape[1], ("utors_slt", "utors_slv")),
        (lambda fl, f2: O.fitors_rem(s, li, *ip, fl, f2),
         O.fitors(s, li, *ip), li.shape[1], ("utors_slt", "utors_slv")),
    ]
    assert not np.isin(np.abs(np.rint(pt[:, 1])), (12, 14)).any()
    for rem, plain, n, keys in cases:
        fl = _rem_flags(n, 4)
        o = rem(fl, 1.0)
        assert np.array_equal(o["f"], plain["f"])
        for k in keys:
            assert o[k] == plain[k], k
        assert rem(np.zeros(n, np.int32), 0.3)["seg_ene"] == 0.0
        a = rem(np.ones(n, np.int32), 0.3)
        etot = sum(plain[k] for k in keys)
        assert abs(a["seg_ene"] - etot) <= 1e-12 * sum(abs(plain[k]) for k in keys)
        assert np.abs(a["f"] - 0.3 * plain["f"]).max() <= 1e-13 * np.abs(plain["f"]).max()
