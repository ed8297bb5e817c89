"""On-device r-RESPA dynamics against the oracle's restatement of the step
(oracle.md_steps: ORAC's verlet, correc, rattle_verlet and rattle_correc in
mtsmd.f's order, the constraint routines pinned bit for bit against the
reference), driven by the oracle's forces; plus energy conservation, time
reversibility, and identical trajectories eager / graph / batched.
"""
import numpy as np
import pytest

import oracle as O
from conftest import relrms

pytestmark = pytest.mark.gpu
KJ_MOL = 4.184  # kJ/mol per kcal/mol


def _prepare(ctx, s, precision=0, relax=40, seed=11, temperature=300.0):
    from orac_hip import md
    from orac_hip.api import Batch, ResidentSystem
    cp = np.concatenate([s.bonds, s.bends], axis=1) if s.bonds is not None else np.zeros((2, 0), np.int32)
    rs = ResidentSystem(ctx, s, precision=precision, corr_pairs=cp, corr_fudge=np.ones(cp.shape[1]))
    dyn = md.Dynamics(rs, s)
    b = Batch(ctx, [rs])
    if relax:
        md.relax_batch(b, relax)
    dyn.maxwell_boltzmann(temperature, seed)
    return rs, dyn, b


def _oracle_forces(s):
    """slow(x), fast(x) at Cartesian positions x as the resident system and
    Dynamics compose them: direct space (list rebuilt) + PME + ferrf over the
    solute's bonds and bends with fudge 1; harmonic solute terms."""
    from orac_hip import md
    cp = np.concatenate([s.bonds, s.bends], axis=1) if s.bonds is not None else np.zeros((2, 0), np.int32)
    oc = np.asarray(s.oc)
    pairs, k, r0 = md.solute_terms(s)

    def slow(x):
        xp = oc @ x
        s2 = s.with_coords(xp[0], xp[1], xp[2])
        rc = s2.rcut + s2.rtol + s2.rnei
        fa = O.ForceArgs(s2)
        if s2.L > 3 * rc:
            nc = int(s2.L // 3.0)
            st = O.lc_index(nc, nc, nc, rc, s2.co)
            nnlpp = O.lc_list(nc, nc, nc, st, rc, s2.co, fa.xpg, fa.ypg, fa.zpg)
        else:
            nnlpp = O.force(s2, O.all_groups_list(s2.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0,
                            iz=0, fa=fa)["nnlpp1"]
        d = O.force(s2, nnlpp, fa=fa, rinn=0.0, rout=s2.rcut, rtolinn=0.0, rtolout=s2.rtol, rneigh=0.0, nthreads=4)
        f = d["f"] + O.fft_pme(s2, variant="fftw", nthreads=4)["f"]
        if cp.shape[1]:
            f = f + O.ferrf(s2, cp, 1, 1.0)["f"]
        return f

    def fast(x):
        return O.harmonic_forces(x, pairs, k, r0)

    return slow, fast


@pytest.mark.parametrize("sysname,ninner", [("water", 1), ("small_solute", 4)])
def test_md_matches_oracle(gpu_ctx, sysname, ninner, request):
    """Ten outer steps of 1 fs from the same positions and velocities: the
    device trajectory against oracle.md_steps, list rebuilt every step."""
    from orac_hip import md
    s = request.getfixturevalue(sysname)
    rs, dyn, b = _prepare(gpu_ctx, s, relax=20)
    x0, v0 = dyn.state()
    md.run_batch(b, 10, dt=1.0, ninner=ninner, list_every=1, graph=False)
    xg, vg = dyn.state()
    fail = dyn.energies()["fail"]
    b.close()
    rs.close()
    first, lens = md.rigid_waters(s)
    cons = O.constraint_tables(first, lens, s.mass)
    slow, fast = _oracle_forces(s)
    xo, vo, fails = O.md_steps(x0, v0, np.asarray(s.mass, np.float64), 10, 1.0, ninner, slow, fast, cons)
    dx, dv = np.abs(xg - xo).max(), np.abs(vg - vo).max()
    print(sysname, "moved", np.abs(xo - x0).max(), "max |dx|", dx, "max |dv|", dv, "of", np.abs(vo).max())
    assert fail == 0 and fails == 0
    assert np.abs(xo - x0).max() > 1e-3  # it moved
    assert dx <= 1e-9
    assert dv <= 1e-9 * np.abs(vo).max()


RCUTM, RTOLM = 5.0, 0.5  # inner shell of the two-shell tests


def _oracle_forces_shells(s, rcutm=RCUTM, rtolm=RTOLM):
    """slow(x), mid(x), fast(x) of a system with two shells: the outer shell
    (rinn = RCUTM switched in over RTOLM, decide_force on its real-space
    Ewald term) + PME + ferrf; the inner shell (0 .. RCUTM switched out);
    the harmonic solute terms."""
    from orac_hip import md
    cp = np.concatenate([s.bonds, s.bends], axis=1) if s.bonds is not None else np.zeros((2, 0), np.int32)
    oc = np.asarray(s.oc)
    pairs, k, r0 = md.solute_terms(s)
    skip = O.decide_force(rcutm, s.alpha)

    def lists(x):
        xp = oc @ x
        s2 = s.with_coords(xp[0], xp[1], xp[2])
        rc = s2.rcut + s2.rtol + s2.rnei
        fa = O.ForceArgs(s2)
        if s2.L > 3 * rc:
            nc = int(s2.L // 3.0)
            st = O.lc_index(nc, nc, nc, rc, s2.co)
            nnlpp = O.lc_list(nc, nc, nc, st, rc, s2.co, fa.xpg, fa.ypg, fa.zpg)
        else:
            nnlpp = O.force(s2, O.all_groups_list(s2.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0,
                            iz=0, fa=fa)["nnlpp1"]
        return s2, fa, nnlpp

    def slow(x):
        s2, fa, nnlpp = lists(x)
        d = O.force(s2, nnlpp, fa=fa, rinn=rcutm, rout=s2.rcut, rtolinn=rtolm, rtolout=s2.rtol, rneigh=0.0,
                    lskip_ewald=skip, nthreads=4)
        f = d["f"] + O.fft_pme(s2, variant="fftw", nthreads=4)["f"]
        if cp.shape[1]:
            f = f + O.ferrf(s2, cp, 1, 1.0)["f"]
        return f

    def mid(x):
        s2, fa, nnlpp = lists(x)
        return O.force(s2, nnlpp, fa=fa, rinn=0.0, rout=rcutm, rtolinn=0.0, rtolout=rtolm, rneigh=0.0,
                       nthreads=4)["f"]

    def fast(x):
        return O.harmonic_forces(x, pairs, k, r0)

    return slow, mid, fast


@pytest.mark.parametrize("sysname", ["water", "small_solute"])
def test_two_shells_sum_to_one_shell(gpu_ctx, sysname, request):
    """The inner and outer shells' switches are complementary over
    [rcutm, rcutm + rtolm] (force.f's inner and outer switch forms with the
    same radii), so with the outer shell's Ewald term kept the two shells
    give the single shell's forces and energies to rounding."""
    from orac_hip.api import ResidentSystem
    s = request.getfixturevalue(sysname)
    assert not O.decide_force(RCUTM, s.alpha)  # erfc(alpha rcutm) above 1e-3: the outer shell keeps Ewald
    rs = ResidentSystem(gpu_ctx, s, precision=0)
    rs.eval(rebuild_list=True)
    one = rs.results()
    rs.set_shells(RCUTM, RTOLM)
    rs.eval(rebuild_list=True)
    two = rs.results()
    rs.set_shells(0.0, 0.0)
    rs.eval(rebuild_list=True)
    back = rs.results()
    rs.close()
    assert relrms(two["f"], one["f"]) < 1e-12
    for k in ("uconf_slv", "ucoul_slv", "eer"):
        assert abs(two[k] - one[k]) <= 1e-12 * abs(one[k]), k
    assert np.array_equal(back["f"], one["f"])


def test_two_shells_mixed_precision(gpu_ctx, small_solute):
    """Mixed precision with two shells: the shells' sum against the mixed
    single shell to the FP32 pair arithmetic's rounding, and against FP64
    within the mixed-precision budget (5e-6 relative RMS)."""
    from orac_hip.api import ResidentSystem
    s = small_solute
    out = {}
    for prec in (0, 1):
        rs = ResidentSystem(gpu_ctx, s, precision=prec)
        rs.eval(rebuild_list=True)
        out[(prec, 1)] = rs.results()["f"]
        rs.set_shells(RCUTM, RTOLM)
        rs.eval(rebuild_list=True)
        out[(prec, 2)] = rs.results()["f"]
        rs.close()
    assert relrms(out[(1, 2)], out[(1, 1)]) < 1e-6
    assert relrms(out[(1, 2)], out[(0, 2)]) < 5e-6


@pytest.mark.parametrize("sysname,rcutm", [("water", RCUTM), ("small_solute", RCUTM), ("water", 6.5)])
def test_md_shells_match_oracle(gpu_ctx, sysname, rcutm, request):
    """Five outer steps of 2 fs with two shells (inner shell twice per step,
    the fast terms twice per inner step) from the same state: the device
    against oracle.md_steps_shells, list rebuilt every step.  At rcutm 6.5
    on the water box erfc(alpha rcutm) is below 1e-3, so the outer shell
    drops its real-space Ewald term (decide_force)."""
    from orac_hip import md
    s = request.getfixturevalue(sysname)
    if rcutm != RCUTM:
        assert O.decide_force(rcutm, s.alpha)
    rs, dyn, b = _prepare(gpu_ctx, s, relax=20)
    rs.set_shells(rcutm, RTOLM)
    b.close()
    from orac_hip.api import Batch
    b = Batch(gpu_ctx, [rs])
    x0, v0 = dyn.state()
    md.run_batch(b, 5, dt=2.0, ninner=2, nmid=2, list_every=1, graph=False)
    xg, vg = dyn.state()
    fail = dyn.energies()["fail"]
    b.close()
    rs.close()
    first, lens = md.rigid_waters(s)
    cons = O.constraint_tables(first, lens, s.mass)
    slow, mid, fast = _oracle_forces_shells(s, rcutm)
    xo, vo, fails = O.md_steps_shells(x0, v0, np.asarray(s.mass, np.float64), 5, 2.0, 2, 2, slow, mid, fast, cons)
    dx, dv = np.abs(xg - xo).max(), np.abs(vg - vo).max()
    print(sysname, "shells: moved", np.abs(xo - x0).max(), "max |dx|", dx, "max |dv|", dv)
    assert fail == 0 and fails == 0
    assert np.abs(xo - x0).max() > 1e-3
    assert dx <= 1e-9
    assert dv <= 1e-9 * np.abs(vo).max()


def test_md_shells_inner_list_ages(gpu_ctx, small_solute):
    """The inner shell walks an inner list filtered (with the skin) at each
    full-list build: six 2 fs outer steps with the lists rebuilt every third
    step give the trajectory of the lists rebuilt every step, to rounding
    (the list order, hence the summation order, differs between builds)."""
    from orac_hip import md
    from orac_hip.api import Batch
    out = []
    for every in (1, 3):
        rs, dyn, b = _prepare(gpu_ctx, small_solute, relax=20)
        rs.set_shells(RCUTM, RTOLM)
        b.close()
        b = Batch(gpu_ctx, [rs])
        md.run_batch(b, 6, dt=2.0, ninner=2, nmid=2, list_every=every, graph=True)
        out.append(dyn.state())
        assert dyn.energies()["fail"] == 0
        b.close()
        rs.close()
    (x1, v1), (x3, v3) = out
    print("inner list aged 3 steps: max |dx|", np.abs(x3 - x1).max())
    assert np.abs(x3 - x1).max() < 1e-9
    assert np.abs(v3 - v1).max() < 1e-9 * np.abs(v1).max()


@pytest.mark.parametrize("precision", [0, 1])
def test_md_shells_batch_equals_alone(gpu_ctx, small_solute, water, precision):
    """Two-shell dynamics keep a system's trajectory independent of its
    batch: alone and batched with a system of another size, graph replay,
    bit for bit, in both precisions."""
    from orac_hip import md
    from orac_hip.api import Batch
    runs = []
    for with_other in (False, True):
        a, da, ba = _prepare(gpu_ctx, small_solute, precision, relax=20)
        w, dw, bw = _prepare(gpu_ctx, water, precision, relax=20)
        for rs in (a, w):
            rs.set_shells(RCUTM, RTOLM)
        ba.close()
        bw.close()
        b = Batch(gpu_ctx, [a, w] if with_other else [a])
        md.run_batch(b, 6, dt=2.0, ninner=2, nmid=2, list_every=3, graph=True)
        runs.append(da.state())
        assert da.energies()["fail"] == 0
        b.close()
        a.close()
        w.close()
    (x1, v1), (x2, v2) = runs
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)


def _rigid_error(dyn):
    from orac_hip import md
    x, _ = dyn.state()
    first, lens = md.rigid_waters(dyn.s)
    f = first - 1
    err = 0.0
    for k, (i, j) in enumerate(((0, 1), (0, 2), (1, 2))):
        d = np.sqrt(((x[:, f + i] - x[:, f + j]) ** 2).sum(axis=0))
        err = max(err, float(np.abs(d - lens[:, k]).max()))
    return err


@pytest.fixture(scope="module")
def water(water216):
    return water216


def test_md_water_nve(gpu_ctx, water):
    """After equilibration at 300 K by velocity rescaling: 10 ps of NVE at
    1 fs, list every 5 steps.  Total-energy drift per atom below
    0.15 kJ/mol/ns (a linear fit over 200 samples; the fluctuation of the
    total energy leaves the fit an uncertainty near 0.005).  The oracle's own
    integrator, with the list rebuilt every step, drifts 0.04 kJ/mol/ns per
    atom on this box at these settings (group-switched cutoff, PME, and the
    reference's MIM SHAKE), the device 0.06: the bound is the algorithm's,
    with room.  The temperature stays near 300 K, and the rigid distances
    keep the size the reference's MIM SHAKE leaves (second order in the
    step: its quadratic term weighs the O-H constraints with the H-H
    constraint's masses)."""
    from orac_hip import md
    rs, dyn, b = _prepare(gpu_ctx, water)
    md.equilibrate(b, [dyn], 300.0, 300, every=10, dt=1.0, ninner=1, list_every=5, graph=True)
    e = []
    for _ in range(200):
        md.run_batch(b, 50, dt=1.0, ninner=1, list_every=5, graph=True)
        e.append(dyn.energies())
    tot = np.array([x["total"] for x in e]) / md.KCAL_MOL * KJ_MOL
    t_ns = 50e-6 * np.arange(len(tot))
    slope = np.polyfit(t_ns, tot, 1)[0] / water.natoms
    temps = np.array([x["temperature"] for x in e])
    rigid = _rigid_error(dyn)
    print("drift kJ/mol/ns/atom", slope, "E sd kJ/mol", tot.std(), "T", temps.min(), temps.max(), "rigid", rigid)
    assert e[-1]["fail"] == 0
    assert abs(slope) < 0.15
    assert 250.0 < temps.mean() < 350.0
    assert rigid < 1e-4
    b.close()
    rs.close()


def test_md_time_reversible(gpu_ctx, small_solute):
    from orac_hip import md
    rs, dyn, b = _prepare(gpu_ctx, small_solute)
    md.run_batch(b, 2, dt=1.0, ninner=2, list_every=1, graph=False)
    x0, v0 = dyn.state()
    md.run_batch(b, 20, dt=1.0, ninner=2, list_every=1, graph=False)
    x1, v1 = dyn.state()
    assert np.abs(x1 - x0).max() > 1e-2  # it moved
    dyn.set_velocities(-v1)
    md.run_batch(b, 20, dt=1.0, ninner=2, list_every=1, graph=False)
    x2, v2 = dyn.state()
    print("return error", np.abs(x2 - x0).max(), np.abs(-v2 - v0).max())
    # exact reversal up to the MIM SHAKE's convergence test (a relative
    # change of 1e-7 in the multipliers, rattle_verlet.f:77), grown over the
    # 40 steps there and back
    assert np.abs(x2 - x0).max() < 5e-5
    assert np.abs(-v2 - v0).max() < 1e-5
    assert dyn.energies()["fail"] == 0
    b.close()
    rs.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_md_graph_batch_eager_identical(gpu_ctx, small_solute, water, precision):
    from orac_hip import md
    from orac_hip.api import Batch
    a1, d1, b1 = _prepare(gpu_ctx, small_solute, precision, relax=20)
    a2, d2, b2 = _prepare(gpu_ctx, small_solute, precision, relax=20)
    w1, e1, c1 = _prepare(gpu_ctx, water, precision, relax=20)
    # eager, alone
    md.run_batch(b1, 12, dt=2.0, ninner=4, list_every=3, graph=False)
    # graph, batched with another system of a different size
    both = Batch(gpu_ctx, [a2, w1])
    md.run_batch(both, 12, dt=2.0, ninner=4, list_every=3, graph=True)
    xa, va = d1.state()
    xb, vb = d2.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert d1.energies()["fail"] == 0 and d2.energies()["fail"] == 0
    for x in (both, b1, b2, c1, a1, a2, w1):
        x.close()
