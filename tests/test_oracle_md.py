"""The integrator's pieces in the oracle: ORAC's constraint routines
(rattle_verlet.f SHAKE, rattle_correc.f RATTLE; the matrix-inversion method
for molecules of at most mim_lim constraints and the bond-by-bond iteration
beyond) and its kick / drift (correc.f, verlet.f), bit for bit against the
reference compiled in place -- which also pins two of the reference's
quirks, the MIM quadratic term weighted with the last constraint's masses
(rattle_verlet.f:406-407) and the molecule cursor that the iterative branch
never advances (:562).  Plus what the mathematics fixes: exact constraints
for a single bond, RATTLE's orthogonality, momentum conservation, and a
time-reversible step."""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")

R_OH, ANG = 0.9572, np.radians(104.52)
R_HH = 2.0 * R_OH * np.sin(ANG / 2.0)
M_O, M_H = 15.9994, 1.008


def _waters(nw, seed=1, vscale=0.01):
    """nw rigid waters at random places and orientations, velocities made
    orthogonal to the bonds."""
    rng = np.random.default_rng(seed)
    n = 3 * nw
    mass = np.tile([M_O, M_H, M_H], nw)
    first = 1 + 3 * np.arange(nw)
    lens = np.tile([R_OH, R_OH, R_HH], (nw, 1))
    x = np.zeros((3, n))
    for w in range(nw):
        c = rng.uniform(-10.0, 10.0, 3)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        sites = (np.zeros(3), np.array([R_OH, 0.0, 0.0]), R_OH * np.array([np.cos(ANG), np.sin(ANG), 0.0]))
        for k, a in enumerate(sites):
            x[:, 3 * w + k] = c + q @ a
    cons = O.constraint_tables(first, lens, mass)
    v = rng.normal(scale=vscale, size=(3, n))
    v, _, rc = O.rattle_correc(x, v, *cons[:3], mass, cons[3])
    assert rc == 0
    return x, v, mass, cons, rng


def _mixed_molecules(seed=4):
    """Molecules of 1, 2 and 3 constraints: a diatomic, a bent triatomic with
    its two bonds constrained (sharing the middle atom), and a water."""
    rng = np.random.default_rng(seed)
    x = np.zeros((3, 8))
    x[:, 0:2] = np.array([[0.0, 1.1], [0.0, 0.0], [0.0, 0.0]])
    x[:, 2:5] = np.array([[3.0, 4.0, 4.6], [0.0, 0.0, 0.8], [0.0, 0.0, 0.0]])
    sites = (np.zeros(3), np.array([R_OH, 0.0, 0.0]), R_OH * np.array([np.cos(ANG), np.sin(ANG), 0.0]))
    for k, a in enumerate(sites):
        x[:, 5 + k] = np.array([0.0, 5.0, 0.0]) + a
    x += rng.normal(scale=0.01, size=x.shape)
    mass = np.array([12.011, 15.9994, 14.007, 12.011, 1.008, M_O, M_H, M_H])
    cnst = np.array([[1, 3, 4, 6, 6, 7], [2, 4, 5, 7, 8, 8]], np.int32)
    d = x[:, cnst[0] - 1] - x[:, cnst[1] - 1]
    dss = (d * d).sum(axis=0)
    ma, mb = mass[cnst[0] - 1], mass[cnst[1] - 1]
    coeff = 2.0 * (ma + mb) / (ma * mb)
    prot = np.array([1, 1, 2, 2, 3, 3, 4, 5, 6], np.int32)
    v = rng.normal(scale=0.01, size=x.shape)
    return x, v, mass, (cnst, dss, coeff, prot), rng


def _drift(x, v, mass, rng, dt=2.0, noise=1e-3):
    return x + dt * v + rng.normal(scale=noise, size=x.shape)


@needs_ref
@pytest.mark.parametrize("mim_lim", [20, 0])
def test_rattle_verlet_exact(mim_lim):
    """SHAKE after a drift, MIM (mim_lim 20) and bond by bond (mim_lim 0)."""
    for maker in (_waters, _mixed_molecules):
        x0, v, mass, cons, rng = maker(7) if maker is _waters else maker()
        x1 = _drift(x0, v, mass, rng)
        a = O.rattle_verlet(2.0, x1, x0, v, *cons[:3], mass, cons[3], mim_lim=mim_lim)
        b = O.ref_rattle_verlet(2.0, x1, x0, v, *cons[:3], mass, cons[3], mim_lim=mim_lim)
        assert a[3] == 0 and b[3] == 0
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert a[2] == b[2]  # iterations per molecule (dnit)


@needs_ref
@pytest.mark.parametrize("mim_lim", [20, 0])
def test_rattle_correc_exact(mim_lim):
    for maker in (_waters, _mixed_molecules):
        x, v, mass, cons, rng = maker(9) if maker is _waters else maker()
        v = v + rng.normal(scale=0.01, size=v.shape)
        a = O.rattle_correc(x, v, *cons[:3], mass, cons[3], mim_lim=mim_lim)
        b = O.ref_rattle_correc(x, v, *cons[:3], mass, cons[3], mim_lim=mim_lim)
        assert a[2] == 0 and b[2] == 0
        assert np.array_equal(a[0], b[0]) and a[1] == b[1]


@needs_ref
def test_kick_and_drift_exact():
    x, v, mass, cons, rng = _waters(6)
    f = rng.normal(size=x.shape)
    assert np.array_equal(O.correc(v, f, mass, 0.75), O.ref_correc(v, f, mass, 0.75))
    xa, va = O.verlet(x, v, f, mass, 0.75)
    xb, vb = O.ref_verlet(x, v, f, mass, 0.75)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_mim_single_bond_is_exact():
    """One constraint: the MIM iteration converges on the exact length (its
    quadratic term uses the right masses when there is one constraint), the
    move is along the old bond, momentum is kept, and v changes by the move
    over dt."""
    rng = np.random.default_rng(2)
    for _ in range(20):
        x0 = np.array([[0.0, 1.2], [0.0, 0.1], [0.0, -0.2]]) + rng.normal(scale=0.1, size=(3, 2))
        mass = rng.uniform(1.0, 20.0, 2)
        d0 = x0[:, 0] - x0[:, 1]
        dss = np.array([(d0 * d0).sum()])
        cnst = np.array([[1], [2]], np.int32)
        coeff = 2.0 * (mass[0] + mass[1]) / (mass[0] * mass[1])
        v = rng.normal(scale=0.02, size=(3, 2))
        x1 = x0 + 2.0 * v + rng.normal(scale=0.01, size=(3, 2))
        xs, vs, _, rc = O.rattle_verlet(2.0, x1, x0, v, cnst, dss, np.array([coeff]), mass, np.array([1, 1], np.int32))
        assert rc == 0
        d = xs[:, 0] - xs[:, 1]
        assert abs(np.sqrt((d * d).sum()) - np.sqrt(dss[0])) < 1e-6 * np.sqrt(dss[0])
        mv = (xs - x1) * mass
        assert np.abs(mv.sum(axis=1)).max() < 1e-14
        move = (xs - x1)[:, 0]
        assert np.linalg.norm(np.cross(move, d0)) < 1e-12 * max(np.linalg.norm(move), 1e-30) * np.linalg.norm(d0) + 1e-15
        assert np.allclose(vs - v, (xs - x1) / 2.0, rtol=0, atol=1e-14)


def test_rattle_orthogonal_and_momentum():
    """RATTLE (MIM, one linear solve): relative velocities orthogonal to
    every bond to rounding, total momentum unchanged."""
    x, v, mass, (cnst, dss, coeff, prot), rng = _waters(12, seed=5)
    v = v + rng.normal(scale=0.02, size=v.shape)
    vr, _, rc = O.rattle_correc(x, v, cnst, dss, coeff, mass, prot)
    assert rc == 0
    i, j = cnst[0] - 1, cnst[1] - 1
    r = x[:, i] - x[:, j]
    dv = vr[:, i] - vr[:, j]
    assert np.abs((r * dv).sum(axis=0)).max() < 1e-15 * 10
    p0, p1 = (v * mass).sum(axis=1), (vr * mass).sum(axis=1)
    assert np.abs(p1 - p0).max() < 1e-14


def test_shake_water_lengths():
    """MIM SHAKE on waters: the H-H constraint (the last one, whose masses
    the quadratic term uses) is met to rounding; the O-H ones to second order
    in the move (the quadratic term weighs O's move with H's mass).  A
    mistake in the linear part would leave first-order errors."""
    x0, v, mass, cons, rng = _waters(20, seed=8)
    x1 = x0 + 1.0 * v
    xs, _, _, rc = O.rattle_verlet(1.0, x1, x0, v, *cons[:3], mass, cons[3])
    assert rc == 0
    cnst, dss = cons[0], cons[1]
    d = xs[:, cnst[0] - 1] - xs[:, cnst[1] - 1]
    err = np.abs(np.sqrt((d * d).sum(axis=0)) - np.sqrt(dss)).reshape(-1, 3)
    move = np.abs(x1 - x0).max()
    assert err[:, 2].max() < 1e-12
    assert err[:, :2].max() < 10.0 * move * move
    # and far below the first-order size of the move
    assert err[:, :2].max() < 1e-2 * move


def test_md_steps_time_reversible():
    """The restated step is time-reversible: reversing the velocities after
    n steps retraces the path (harmonic fast terms on a chain, a pair force
    between waters as the slow force)."""
    x, v, mass, cons, rng = _waters(8, seed=11, vscale=0.005)
    pairs = np.array([[1, 4, 7], [4, 7, 10]], np.int32)
    k = np.array([0.01, 0.02, 0.015])
    d = x[:, pairs[0] - 1] - x[:, pairs[1] - 1]
    r0 = np.sqrt((d * d).sum(axis=0)) * 0.97

    def fast(y):
        return O.harmonic_forces(y, pairs, k, r0)

    def slow(y):
        # soft pairwise repulsion between all oxygen sites
        f = np.zeros_like(y)
        oxy = np.arange(0, y.shape[1], 3)
        for a in oxy:
            for b in oxy:
                if a < b:
                    dd = y[:, a] - y[:, b]
                    r2 = (dd * dd).sum()
                    c = 1e-3 / (r2 * r2)
                    f[:, a] += c * dd
                    f[:, b] -= c * dd
        return f

    xa, va, fails = O.md_steps(x, v, mass, 15, 1.0, 3, slow, fast, cons)
    assert fails == 0 and np.abs(xa - x).max() > 1e-2
    xb, vb, fails = O.md_steps(xa, -va, mass, 15, 1.0, 3, slow, fast, cons)
    assert fails == 0
    # MIM's tolerance (1e-7 relative on the multipliers) bounds the return
    assert np.abs(xb - x).max() < 1e-6
    assert np.abs(-vb - v).max() < 1e-6


def test_md_steps_end_state_invariants():
    """What mtsmd.f's order fixes at the end of an outer step, whatever the
    forces: the last call is a RATTLE, so the relative velocities of every
    constrained pair are orthogonal to its bond (to rounding); the last
    drift was SHAKEn, so the bond lengths hold to the MIM method's second
    order.  A step that ended on a kick, or skipped a SHAKE, fails one."""
    x, v, mass, cons, rng = _waters(10, seed=21, vscale=0.01)
    cnst, dss = cons[0], cons[1]

    def slow(y):
        return 1e-3 * np.sin(y)  # any smooth force field, of the size of the pair forces above

    def fast(y):
        return 3e-4 * np.cos(2.0 * y)

    xa, va, fails = O.md_steps(x, v, mass, 4, 2.0, 3, slow, fast, cons)
    assert fails == 0
    i, j = cnst[0] - 1, cnst[1] - 1
    r = xa[:, i] - xa[:, j]
    dv = va[:, i] - va[:, j]
    assert np.abs((r * dv).sum(axis=0)).max() < 1e-14
    lens = np.sqrt((r * r).sum(axis=0))
    assert np.abs(lens - np.sqrt(dss)).max() < 1e-4
    # and the same with two shells
    xb, vb, fails = O.md_steps_shells(x, v, mass, 2, 4.0, 2, 3, slow, lambda y: 0.5 * slow(y), fast, cons)
    assert fails == 0
    r = xb[:, i] - xb[:, j]
    dv = vb[:, i] - vb[:, j]
    assert np.abs((r * dv).sum(axis=0)).max() < 1e-14
