"""Batched evaluation: several resident systems, one kernel launch per stage.

Each system's result in a batch must be the one it gets alone (bit for bit
with the same tuning: every reduction is per system and in fixed order),
HREM replicas must each match the FP64 oracle, graph replays must match the
eager evaluation on every replay, and the autotuner must not change forces.
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


def _resident(ctx, s, precision=0):
    from orac_hip.api import ResidentSystem
    cp, cf = _corr(s)
    return ResidentSystem(ctx, s, precision=precision, corr_pairs=cp, corr_fudge=cf)


@pytest.fixture(scope="module")
def mixed_sizes():
    from orac_hip import synth
    return [synth.config("water216"), synth.small_solvated(rcut=7.0),
            synth.water_box(seed=77, L=24.0, nside=8, nfft=(24, 30, 25), rcut=9.0)]


def _same(a, b):
    assert np.array_equal(a["f"], b["f"])
    for k in ("eer", "uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss", "corr_slt",
              "corr_slv"):
        assert a[k] == b[k], k
    assert np.array_equal(a["stress"], b["stress"])
    assert a["list_entries"] == b["list_entries"]


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_alone(gpu_ctx, mixed_sizes, precision):
    from orac_hip.api import Batch
    systems = [_resident(gpu_ctx, s, precision) for s in mixed_sizes]
    alone = []
    for rs in systems:
        rs.eval(rebuild_list=True)
        alone.append(rs.results())
    b = Batch(gpu_ctx, systems)
    b.eval(rebuild_list=True)
    for r, a in zip(b.results(), alone):
        _same(r, a)
    # order within the batch does not matter either
    b2 = Batch(gpu_ctx, systems[::-1])
    b2.eval(rebuild_list=True)
    for r, a in zip(b2.results(), alone[::-1]):
        _same(r, a)
    for x in (b, b2, *systems):
        x.close()


def test_batch_replicas_match_oracle(gpu_ctx):
    from orac_hip import synth
    from orac_hip.api import Batch
    from test_gpu_resident import oracle_total
    reps = synth.hrem_replicas(n=4, s_min=0.3, n_solute=40, radius=5.0, n_water=380, L=24.0, nfft=(24, 24, 24),
                               rcut=7.0)
    systems = [_resident(gpu_ctx, s, 0) for s in reps]
    b = Batch(gpu_ctx, systems)
    b.eval(rebuild_list=True)
    res = b.results()
    for s, g in zip(reps, res):
        cp, cf = _corr(s)
        o = oracle_total(s, cp, cf)
        assert relrms(g["f"], o["f"]) < 1e-9
        assert rel(g["eer"], o["eer"]) < 1e-10
        assert rel(g["ucoul_slt"], o["d"]["ucoul_slt"]) < 1e-10
    # the replicas differ (scaled solute), so a mix-up between them would show
    assert rel(res[0]["ucoul_slt"], res[-1]["ucoul_slt"]) > 0.1
    b.close()
    for x in systems:
        x.close()


def test_batch_graph_replays(gpu_ctx, mixed_sizes):
    from orac_hip.api import Batch
    systems = [_resident(gpu_ctx, s, 0) for s in mixed_sizes]
    b = Batch(gpu_ctx, systems)
    b.eval(rebuild_list=True)
    ref = b.results()
    for k in range(4):
        # eager work and synchronisation between replays
        systems[0].eval(rebuild_list=True)
        gpu_ctx.sync()
        b.eval(rebuild_list=(k % 2 == 0), graph=True)
        for r, a in zip(b.results(), ref):
            _same(r, a)
    b.close()
    for x in systems:
        x.close()


def test_autotune_keeps_forces(gpu_ctx, mixed_sizes):
    from orac_hip.api import Batch
    systems = [_resident(gpu_ctx, s, 1) for s in mixed_sizes]
    b = Batch(gpu_ctx, systems)
    b.eval(rebuild_list=True)
    ref = b.results()
    tun, ms = b.autotune(reps=2)
    assert ms > 0 and tun["lds_elems"] in (256, 512, 1024) and tun["pair_waves"] in (0, 2, 3, 4)
    # the log holds every timed candidate, the kept one the fastest of them
    log = b.tune_log()
    assert len(log) > 6 and {t["persistent"] for t, _ in log} == {0, 1}
    assert min(m for _, m in log) == ms and (tun, ms) in log
    for t in ({"lds_elems": 256, "pair_waves": 2, "spread_tiled": 0, "xcd_remap": 0, "cell_order": 1, "persistent": 1},
              {"lds_elems": 1024, "pair_waves": 4, "spread_tiled": 1, "xcd_remap": 1, "cell_order": 0, "persistent": 0},
              {"lds_elems": 512, "pair_waves": 3, "spread_tiled": 2, "xcd_remap": 1, "cell_order": 1, "persistent": 1},
              {"lds_elems": 512, "pair_waves": 4, "spread_tiled": 1, "xcd_remap": 1, "cell_order": 0, "persistent": 0},
              tun):
        b.set_tuning(**t)
        b.eval(rebuild_list=True)
        for r, a in zip(b.results(), ref):
            assert np.array_equal(r["f"], a["f"])
            # FFT tiling and owner order regroup the per-block partial sums only
            assert rel(r["eer"], a["eer"]) < 1e-12
            assert rel(r["ucoul_slv"], a["ucoul_slv"]) < 1e-12
    b.close()
    for x in systems:
        x.close()


def test_batch_rejects_mixed_precision(gpu_ctx, mixed_sizes):
    from orac_hip.api import Batch, OracHipError
    a = _resident(gpu_ctx, mixed_sizes[0], 0)
    c = _resident(gpu_ctx, mixed_sizes[0], 1)
    with pytest.raises(OracHipError) as e:
        Batch(gpu_ctx, [a, c])
    assert e.value.code == 1
    a.close()
    c.close()
