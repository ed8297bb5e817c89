"""force_rem (force_REM.f) on the GPU against the oracle's restatement, which
is pinned bit for bit to the reference compiled in place: the three
seg_select kinds (rem.f90:631-645), the single and switched double shell,
the charges-off lskip_force + ljintra case, and the error for the branches
that read locals left by an earlier pair.  Tolerances are the FP64 plain
path's: forces relative RMS <= 1e-9, energies and seg_ene <= 1e-10."""
import numpy as np
import pytest

import oracle as O
from conftest import rel, relrms
from test_gpu_parity import SHELLS, _oracle_list, _stress_close

pytestmark = pytest.mark.gpu

F_TOL = 1e-9
E_TOL = 1e-10
E6 = ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss")


def _segment(s, seed=3):
    """The solute's first half plus five waters."""
    rng = np.random.default_rng(seed)
    rem = np.zeros(s.natoms, np.int32)
    nsol = int((np.asarray(s.ss_index) == 1).sum())
    rem[: nsol // 2] = 1
    rem[nsol + 3 * rng.choice((s.natoms - nsol) // 3, 5, replace=False)] = 1
    return rem


def _force_rem_gpu(ctx, s, nnlpp, fa, rem, kind, f3, *, rinn, rout, rtolinn, rtolout, rneigh, lskip_force=False,
                   ljintra=False, lskip_ewald=False, ewald=True):
    n = s.natoms
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    st = np.zeros((3, 3), order="F")
    nn1 = np.zeros(len(nnlpp) + s.ngrp + 16, np.int32)
    r = ctx.force(s.ss_index, s.xp0, s.yp0, s.zp0, fa.xpg, fa.ypg, fa.zpg, s.charge, s.nbtype, s.type_table,
                  s.ntypes, n, s.atomg, fa.xpcm, fa.ypcm, fa.zpcm, None, s.atomp, s.co, s.ecc12, s.ecc6, s.alpha,
                  fa.mapnl, s.ngrp, s.grppt, fpx, fpy, fpz, st, nnlpp, nn1, rneigh, rinn, rout, rtolinn, rtolout,
                  s.pmass, 1, s.mass, 1e20, lskip_ewald, ewald, rem=(rem, kind, f3, lskip_force, ljintra))
    r.update(f=np.stack([fpx, fpy, fpz]), stress=st, nnlpp1=nn1[:r["ncount1"]])
    return r


def _check(g, o):
    assert relrms(g["f"], o["f"]) < F_TOL
    for k in E6:
        if o[k] == 0.0:
            assert g[k] == 0.0, k
        else:
            assert rel(g[k], o[k]) < E_TOL, (k, g[k], o[k])
    assert _stress_close(g["stress"], o["stress"])
    assert abs(g["seg_ene"] - o["seg_ene"]) <= E_TOL * max(abs(o["seg_ene"]), sum(abs(o[k]) for k in E6) * 1e-3)
    assert g["ncount1"] == o["ncount1"] and np.array_equal(g["nnlpp1"], o["nnlpp1"])


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("shell", [0, 1])
def test_force_rem_matches_oracle(gpu_ctx, small_solute, kind, shell):
    s = small_solute
    nnlpp, fa, _ = _oracle_list(s)
    rem = _segment(s)
    sh = dict(SHELLS[shell])
    if sh["rout"] is None:
        sh["rout"] = s.rcut
    g = _force_rem_gpu(gpu_ctx, s, nnlpp, fa, rem, kind, 0.55, **sh)
    o = O.force_rem(s, nnlpp, rem, rem_kind=kind, rem_f3=0.55, fa=fa, **sh)
    print("kind", kind, "shell", shell, "f relrms", relrms(g["f"], o["f"]), "seg", g["seg_ene"], o["seg_ene"])
    assert o["seg_ene"] != 0.0
    _check(g, o)


def test_force_rem_charges_off(gpu_ctx, small_solute):
    """lskip_force with ljintra: the charges count as zero; without ljintra
    nothing is computed and the caller's arrays are left as they are."""
    s = small_solute
    nnlpp, fa, _ = _oracle_list(s)
    rem = _segment(s)
    sh = dict(rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3, rneigh=0.0)
    g = _force_rem_gpu(gpu_ctx, s, nnlpp, fa, rem, 0, 0.7, lskip_force=True, ljintra=True, **sh)
    o = O.force_rem(s, nnlpp, rem, rem_kind=0, rem_f3=0.7, fa=fa, lskip_force=True, ljintra=True, **sh)
    _check(g, o)
    assert g["ucoul_slt"] == 0.0
    g = _force_rem_gpu(gpu_ctx, s, nnlpp, fa, rem, 0, 0.7, lskip_force=True, ljintra=False, **sh)
    assert not g["f"].any() and g["seg_ene"] == 0.0 and g["ncount1"] == 0


@pytest.mark.parametrize("lskip,ewald", [(True, True), (True, False), (False, False)])
def test_force_rem_order_dependent_branches_are_refused(gpu_ctx, small_solute, lskip, ewald):
    from orac_hip.api import OracHipError
    s = small_solute
    nnlpp, fa, _ = _oracle_list(s)
    with pytest.raises(OracHipError) as e:
        _force_rem_gpu(gpu_ctx, s, nnlpp, fa, _segment(s), 0, 0.5, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.3,
                       rneigh=0.0, lskip_ewald=lskip, ewald=ewald)
    assert e.value.code == 4


def test_force_rem_unit_factor_is_force(gpu_ctx, small_solute):
    """rem_f3 = 1 on the GPU gives the plain force entry point's forces (to
    rounding: the REM pairs take the full-list kernel, the plain call on
    groups of at most three atoms the owner-atom loop form, whose summation
    order and reciprocal refinements differ)."""
    from test_gpu_parity import _force_gpu
    s = small_solute
    nnlpp, fa, _ = _oracle_list(s)
    sh = dict(rinn=4.0, rout=7.0, rtolinn=0.5, rtolout=0.5, rneigh=0.6)
    g = _force_rem_gpu(gpu_ctx, s, nnlpp, fa, _segment(s), 1, 1.0, **sh)
    p = _force_gpu(gpu_ctx, s, nnlpp, fa, **sh)
    assert relrms(g["f"], p["f"]) < 1e-11
    for k in E6:
        assert rel(g[k], p[k]) < 1e-12, k


# ---------------------------------------------------------------- bonded REM twins (FKBP12 topology)

@pytest.fixture(scope="module")
def fk():
    from orac_hip import formats
    s, _ = formats.fkbp12()
    return s, np.load(formats.FKBP12_PATH)


def _flags(n, seed):
    return (np.random.default_rng(seed).random(n) < 0.4).astype(np.int32)


@pytest.mark.parametrize("ewald", [True, False])
def test_fnb14_rem_matches_oracle(gpu_ctx, fk, ewald):
    s, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    i14 = np.asfortranarray(g["int14"].astype(np.int32))
    fl = _flags(i14.shape[1], 1)
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    e = gpu_ctx.fnb14_rem(s.ss_index, x, y, z, s.charge, n, s.ecc12, s.ecc6, True, ewald, s.alpha, i14, i14.shape[1],
                          g["type14"], 0.83333, 0.5, fpx, fpy, fpz, fl, 0.61)
    o = O.fnb14_rem(s, i14, g["type14"], 0.83333, 0.5, fl, 0.61, ewald=ewald, cut=True)
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    for got, k in zip(e, ("uconf_slt", "uconf_slv", "ucoul_slt", "ucoul_slv", "seg_ene")):
        assert (got == 0.0 and o[k] == 0.0) or rel(got, o[k]) < 1e-12, k
    # no pair: the incoming seg_ene is left as it was (fnb14_rem.f:100)
    assert gpu_ctx.fnb14_rem(s.ss_index, x, y, z, s.charge, n, s.ecc12, s.ecc6, True, ewald, s.alpha, i14, 0,
                             g["type14"], 0.83333, 0.5, fpx, fpy, fpz, fl, 0.61, seg_ene=7.5)[4] == 7.5


@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fptors_rem_matches_oracle(gpu_ctx, fk, which):
    """The topology's torsions, and 1,700 torsions cycling through every
    multiplicity of both signs, including the 12 and 14 forms whose REM
    evaluation differs from fptors'."""
    s, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    lt, pt = g["ltor"].astype(np.int32), g["potto"]
    pa, pb = pt[:, 1], pt[:, 0]
    if which == "branch":
        mult = np.array([1, 2, 3, 4, 5, 6, 7, 8, 12, 14, -1, -2, -3, -4, -6, -12, -14], float)
        lt = lt[:, :1700]
        pa, pb = mult[np.arange(1700) % len(mult)], pt[:1700, 0]
    lt = np.asfortranarray(lt)
    fl = _flags(lt.shape[1], 2)
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b, seg = gpu_ctx.fptors_rem(s.ss_index, lt, lt.shape[1], x, y, z, pa, pb, fpx, fpy, fpz, n, fl, 0.47)
    o = O.fptors_rem(s, lt, pa, pb, fl, 0.47)
    assert o["status"] == 0
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    assert rel(a, o["utors_slt"]) < 1e-12 and b == o["utors_slv"] == 0.0
    assert rel(seg, o["seg_ene"]) < 1e-12


@pytest.mark.parametrize("which", ["topology", "branch"])
def test_fitors_rem_matches_oracle(gpu_ctx, fk, which):
    s, g = fk
    x, y, z = s.cartesian()
    n = s.natoms
    if which == "topology":
        lt, pots = g["litr"].astype(np.int32), g["itor_pots"]
    else:
        lt, pots = g["ltor"].astype(np.int32)[:, :800], g["itor_branch_pots"]
    lt = np.asfortranarray(lt)
    fl = _flags(lt.shape[1], 3)
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    a, b, seg = gpu_ctx.fitors_rem(s.ss_index, lt, lt.shape[1], x, y, z, *pots, fpx, fpy, fpz, n, fl, 0.52)
    o = O.fitors_rem(s, lt, *pots, fl, 0.52)
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
    assert rel(a, o["utors_slt"]) < 1e-12 and b == o["utors_slv"] == 0.0
    assert rel(seg, o["seg_ene"]) < 1e-12
