"""The batched and serving workloads at their full sizes on the GPU.

* 32 HREM replicas of the 24,460-atom solvated solute (solute charges
  scaled by sqrt(s), solute LJ epsilon by s, s geometric 1 -> 0.3), mixed
  precision, one batch: every replica's forces, energies and virial the same
  in the batch as alone (bit for bit), and replicas 0 and 31 against the FP64
  oracle (forces relative RMS <= 5e-6, the mixed-precision bar).
* 16 lambda-windows of a 48-atom ligand decoupled in a 40 A water box
  (6,399 atoms, PME 48^3, order 5, cutoff 9 A), one batch: every window the
  same as alone, window 0 and window 15 against the oracle's alchemical step
  (FP64: forces <= 1e-9, dwrk <= 1e-10 of the energy scale).
* a served request of more than 100,000 atoms (a 20,000-atom lattice solute
  in water): its final state the same alone and in a mix with two smaller
  requests that arrive while it runs.
"""
import numpy as np
import pytest

from conftest import rel, relrms

pytestmark = pytest.mark.gpu


def _corr(s):
    if s.bonds is None:
        return np.zeros((2, 0), np.int32), np.zeros(0)
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    return pairs, np.ones(pairs.shape[1])


def _resident(ctx, s, precision):
    from orac_hip.api import ResidentSystem
    cp, cf = _corr(s)
    return ResidentSystem(ctx, s, precision=precision, corr_pairs=cp, corr_fudge=cf)


def _same(a, b):
    assert np.array_equal(a["f"], b["f"])
    for k in ("eer", "uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss", "corr_slt",
              "corr_slv"):
        assert a[k] == b[k], k
    assert np.array_equal(a["stress"], b["stress"])


def test_hrem32_mixed_batch_full_size(gpu_ctx):
    from orac_hip import synth
    from orac_hip.api import Batch
    from test_gpu_resident import oracle_total
    reps = synth.hrem_replicas(n=32)
    assert all(s.natoms == 24460 for s in reps)
    systems = [_resident(gpu_ctx, s, 1) for s in reps]
    alone = []
    for k, rs in enumerate(systems):
        rs.eval(rebuild_list=True)
        alone.append(rs.results())
    print("alone done", flush=True)
    b = Batch(gpu_ctx, systems)
    b.eval(rebuild_list=True)
    res = b.results()
    for k, (r, a) in enumerate(zip(res, alone)):
        _same(r, a)
    print("batched == alone for 32 replicas", flush=True)
    for k in (0, 31):
        cp, cf = _corr(reps[k])
        o = oracle_total(reps[k], cp, cf, nthreads=16)
        err = relrms(res[k]["f"], o["f"])
        print("replica", k, "mixed f relrms vs FP64 oracle", err, flush=True)
        assert err <= 5e-6
        assert rel(res[k]["eer"], o["eer"]) < 1e-5
    # the replicas differ (scaled solute): a mix-up would show
    assert rel(res[0]["ucoul_slt"], res[31]["ucoul_slt"]) > 0.1
    b.close()
    for x in systems:
        x.close()


def test_lambda16_windows_full_size(gpu_ctx):
    from orac_hip import synth
    from orac_hip.api import Batch, ResidentSystem
    from test_gpu_alchemy import oracle_alchemy_total
    s, al, states = synth.alchemy_windows(n_windows=16)
    assert s.natoms == 6399 and tuple(s.nfft) == (48, 48, 48)
    cp, cf = _corr(s)

    def window(lam):
        rs = ResidentSystem(gpu_ctx, s, precision=0, corr_pairs=cp, corr_fudge=cf)
        rs.set_alchemy(al, lam)
        return rs

    alone = []
    for lam in states:
        rs = window(lam)
        rs.eval(rebuild_list=True)
        alone.append((rs.results(), rs.alchemy_results()))
        rs.close()
    rss = [window(lam) for lam in states]
    b = Batch(gpu_ctx, rss)
    b.eval(rebuild_list=True)
    for w, (rs, (r0, a0)) in enumerate(zip(rss, alone)):
        r, a = rs.results(), rs.alchemy_results()
        assert np.array_equal(r["f"], r0["f"]), w
        assert abs(a["dwrk"] - a0["dwrk"]) <= 1e-12 * max(1.0, abs(a0["dwrk"])), w
    print("16 windows batched == alone", flush=True)
    for w in (0, 15):
        o = oracle_alchemy_total(s, al, states[w], cp, cf, synth)
        r, a = rss[w].results(), rss[w].alchemy_results()
        scale = abs(o["p"]["eer"]) + sum(abs(o["d"][k]) for k in
                                         ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"))
        err = relrms(r["f"], o["f"])
        print("window", w, "f relrms", err, "dwrk", a["dwrk"], o["dwrk"], flush=True)
        assert err < 1e-9
        assert rel(r["eer"], o["p"]["eer"]) < 1e-10
        assert abs(a["dwrk"] - o["dwrk"]) <= 1e-10 * scale
    b.close()
    for rs in rss:
        rs.close()


def test_served_100k_request_alone_equals_mix(gpu_ctx):
    import dataclasses

    from orac_hip import serve
    big = serve.serving_request(seed=901, n_solute=20000, rcut=9.5)
    assert big.natoms > 100000
    small = [serve.serving_request(seed=902 + k, n_solute=n, rcut=9.0) for k, n in enumerate((300, 2000))]
    reqs = [serve.Request(rid=0, system=big, nsteps=12, seed=11),
            serve.Request(rid=1, system=small[0], nsteps=20, seed=12),
            serve.Request(rid=2, system=small[1], nsteps=8, seed=13)]
    serve.relax_requests(gpu_ctx, reqs, niter=100)
    print("relaxed", [r.system.natoms for r in reqs], flush=True)

    def fresh(arrivals):
        return [dataclasses.replace(r, arrival=arrivals[i], admitted=None, finished=None, final_positions=None,
                                    final_velocities=None, energies=None) for i, r in enumerate(reqs)]

    alone = fresh([0.0, 0.0, 0.0])
    for r in alone:
        serve.Server(gpu_ctx, hbm_budget=1 << 38, chunk=4).run([r])
    mixed = fresh([0.0, 0.005, 0.01])
    stats = serve.Server(gpu_ctx, hbm_budget=1 << 38, chunk=4).run(mixed)
    print(stats, flush=True)
    assert stats["max_active"] >= 2
    for a, m in zip(alone, mixed):
        assert m.finished is not None and m.energies["fail"] == 0
        assert np.array_equal(a.final_positions, m.final_positions), a.rid
        assert np.array_equal(a.final_velocities, m.final_velocities), a.rid
