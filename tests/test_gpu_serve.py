"""Serving: continuous batching of MD segment requests.

Each request's final state must be the same alone and in the mix (bit for
bit: systems are independent in a batch, keep their own list schedules, and
running systems keep their forces when the batch changes), every request
must finish, and admission must respect the HBM budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def requests(gpu_ctx):
    from orac_hip import serve
    reqs = []
    for k, (n, steps) in enumerate([(60, 24), (150, 16), (100, 30), (40, 20)]):
        s = serve.serving_request(seed=100 + k, n_solute=n, rcut=9.0)
        reqs.append(serve.Request(rid=k, system=s, nsteps=steps, arrival=0.0, seed=7 + k))
    serve.relax_requests(gpu_ctx, reqs, niter=150)
    return reqs


def _fresh(reqs, arrivals=None):
    import dataclasses
    return [dataclasses.replace(r, arrival=(arrivals[i] if arrivals is not None else r.arrival), admitted=None,
                                finished=None, final_positions=None, final_velocities=None, energies=None)
            for i, r in enumerate(reqs)]


def test_serve_alone_equals_mix(gpu_ctx, requests):
    from orac_hip import serve
    alone = []
    for r in _fresh(requests):
        srv = serve.Server(gpu_ctx, hbm_budget=1 << 36, chunk=5)
        srv.run([r])
        alone.append(r)
    mixed = _fresh(requests, arrivals=[0.0, 0.0, 0.01, 0.02])
    srv = serve.Server(gpu_ctx, hbm_budget=1 << 36, chunk=5)
    stats = srv.run(mixed)
    print(stats)
    assert stats["max_active"] >= 2
    for a, m in zip(alone, mixed):
        assert m.finished is not None and m.energies["fail"] == 0
        assert np.array_equal(a.final_positions, m.final_positions), a.rid
        assert np.array_equal(a.final_velocities, m.final_velocities), a.rid
        assert np.isfinite(m.energies["total"])


def test_serve_admission_budget(gpu_ctx, requests):
    from orac_hip import serve
    reqs = _fresh(requests, arrivals=[0.0, 0.0, 0.0, 0.0])
    est = sorted(serve.estimate_bytes(r.system) for r in reqs)
    budget = sum(est[:3]) - 1  # requests are charged at least their estimate: never three at once
    srv = serve.Server(gpu_ctx, hbm_budget=budget, chunk=4)
    stats = srv.run(reqs)
    assert all(r.finished is not None for r in reqs)
    assert stats["max_active"] <= 2
    # arrival order is kept at admission
    adm = [r.admitted for r in sorted(reqs, key=lambda r: r.rid)]
    assert adm == sorted(adm)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
