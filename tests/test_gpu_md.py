"""On-device r-RESPA dynamics: constraints, energy conservation, time
reversibility, and identical trajectories eager / graph / batched.

The integrator has no CPU oracle of the reference's own; it is pinned by the
invariants of NVE velocity Verlet with SHAKE/RATTLE: rigid distances held,
total energy conserved, the trajectory retraced when velocities are reversed,
and the same bits whatever the launch form.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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
    from orac_hip import md
    rs, dyn, b = _prepare(gpu_ctx, water)
    md.run_batch(b, 5, dt=1.0, ninner=1, list_every=5, graph=False)
    e0 = dyn.energies()
    md.run_batch(b, 200, dt=1.0, ninner=1, list_every=5, graph=False)
    e1 = dyn.energies()
    print("E0", e0, "E1", e1)
    assert e1["fail"] == 0
    assert _rigid_error(dyn) < 1e-9
    # the jittered lattice is far from equilibrium: potential energy turns
    # into heat, so only a sane range is asked of the temperature
    assert 50.0 < e1["temperature"] < 3000.0
    # NVE: total-energy change small against the kinetic energy
    assert abs(e1["total"] - e0["total"]) < 0.01 * e0["kinetic"]
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
    assert np.abs(x2 - x0).max() < 1e-6
    assert np.abs(-v2 - v0).max() < 1e-6
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
