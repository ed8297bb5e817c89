"""Throughput of the ORAC nonbonded force evaluation on one MI355X.

Workload: the 24,460-atom solvated-solute system (61 A box, 1,660-atom
solute with chain exclusions + 7,600 rigid waters; alpha 0.35, PME grid 64^3,
order 5, cutoff 10 A switched over 0.3 A), synthetic coordinates from the
seeded generator.  One step = one full force evaluation with all state in
HBM: group and molecule centres, linked-cell list rebuild, direct-space
Ewald-erfc + LJ pairs, SPME (spread, r2c FFT, influence function, c2r FFT,
gather) and the per-step exclusion corrections.

Prints one JSON line.  ``value`` is FP64 force evaluations per second of
the single system (the plain path).  Beside it: the same system in mixed
precision, and ``batched``: 32 HREM replicas of the system (solute charges
scaled by sqrt(s), solute LJ epsilon by s, s geometric 1 -> 0.3, independent
water placements) in mixed precision, one kernel launch per stage for all
replicas, autotuned, replayed as a HIP graph; its aggregate evaluations per
second against 32 sequential plain FP64 evaluations.  ``roofline`` is the
FP64 pair kernel, timed with HIP events around that kernel on the context
stream.  ``cpu_baseline`` times the FP64 CPU oracle (OpenMP C restatement
of the reference routines) on this host on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "orac-porting_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FLOP_PER_PAIR = 89


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def evaluated_pairs(s):
    """Atom pairs the direct kernel evaluates: group pairs with centre distance
    below rcut + rtol (both switched and unswitched), times their atom counts,
    minus excluded inter-group pairs inside that range."""
    from scipy.spatial import cKDTree
    L = s.L
    xpg = s.group_centres()
    cart = (xpg.T * (L / 2.0)) % L
    rout0 = s.rcut + s.rtol
    tree = cKDTree(cart, boxsize=L)
    pairs = tree.query_pairs(rout0, output_type="ndarray")
    size = (s.grppt[1] - s.grppt[0] + 1).astype(np.int64)
    total = int(np.sum(size[pairs[:, 0]] * size[pairs[:, 1]]))
    ex = s.exclusion_pairs()
    if ex.size:
        ga, gb = s.atomg[ex[0] - 1] - 1, s.atomg[ex[1] - 1] - 1
        inter = ga != gb
        d = cart[ga[inter]] - cart[gb[inter]]
        d -= L * np.round(d / L)
        total -= int(np.sum(np.sum(d * d, axis=1) < rout0 * rout0))
    return total


def corrections(s):
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    return pairs, np.ones(pairs.shape[1])


def resident(ctx, s, precision):
    from orac_hip.api import ResidentSystem
    cp, cf = corrections(s)
    return ResidentSystem(ctx, s, precision=precision, corr_pairs=cp, corr_fudge=cf)


def profile_stages(obj, nprof):
    """Mean stage milliseconds over nprof profiled evaluations (HIP events)."""
    from orac_hip.api import STAGES
    obj.profile(True)
    st = np.zeros(len(STAGES))
    for _ in range(nprof):
        obj.eval(rebuild_list=True)
        st += np.array(list(obj.stage_ms().values()))
    obj.profile(False)
    return dict(zip(STAGES, st / nprof))


def time_single(ctx, s, precision, steps, warmup):
    rs = resident(ctx, s, precision)
    for _ in range(warmup):
        rs.eval(rebuild_list=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        rs.eval(rebuild_list=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    res = rs.results()
    stages = profile_stages(rs, max(5, min(steps, 20)))
    rs.close()
    return dt, res, stages


def time_batched(ctx, reps, precision, steps, warmup, tune=True):
    from orac_hip.api import Batch
    systems = [resident(ctx, s, precision) for s in reps]
    b = Batch(ctx, systems)
    tuning, tune_ms = (b.autotune(reps=3) if tune else (b.tuning(), None))
    for _ in range(warmup):
        b.eval(rebuild_list=True, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.eval(rebuild_list=True, graph=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    res0 = systems[0].results()
    stages = profile_stages(b, 3)
    mem = sum(x.device_bytes() for x in systems)
    b.close()
    for x in systems:
        x.close()
    return dict(dt=dt, res0=res0, stages=stages, tuning=tuning, tune_ms=tune_ms, device_bytes=mem)


def cpu_baseline(s, corr, nthreads):
    import oracle as O
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nc = int(s.L // 3.0)
    st = O.lc_index(nc, nc, nc, rc, s.co)
    nsample = 0
    t0 = time.perf_counter()
    f = None
    while True:
        nnlpp = O.lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
        d = O.force(s, nnlpp, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0,
                    nthreads=nthreads)
        p = O.fft_pme(s, variant="fftw", nthreads=nthreads)
        c = O.ferrf(s, corr[0], 1, 1.0) if corr[0].shape[1] else None
        f = d["f"] + p["f"] + (c["f"] if c else 0.0)
        nsample += 1
        if time.perf_counter() - t0 > 10.0 or nsample >= 20:
            break
    dt = time.perf_counter() - t0
    return nsample / dt, nsample, dict(f=f, eer=p["eer"])


def relrms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--replicas", type=int, default=32, help="HREM replicas of the batched leg (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device copy of this size first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (one system or one batch per device); --gpus {args.gpus} is not "
              "supported: sharding across GPUs is not part of this build.", file=sys.stderr)
        sys.exit(2)

    from orac_hip import synth
    from orac_hip.api import Context

    s = synth.config("solvated24k")
    corr = corrections(s)
    npairs = evaluated_pairs(s)
    log(f"system: {s.natoms} atoms, {s.ngrp} groups, {npairs} evaluated pairs")

    with Context(0) as ctx:
        dev = ctx.device_name()
        if args.copy_probe_mb:
            import torch
            n = args.copy_probe_mb * (1 << 20) // 8
            a = torch.ones(n, dtype=torch.float64, device="cuda")
            b = torch.empty_like(a)
            b.copy_(a)
            torch.cuda.synchronize()
            del a, b
            log(f"copy probe: {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written")
        dt64, r64, st64 = time_single(ctx, s, 0, args.steps, args.warmup)
        log(f"fp64 single: {1e3 * dt64 / args.steps:.4f} ms/eval")
        dt32, r32, st32 = time_single(ctx, s, 1, args.steps, args.warmup)
        log(f"mixed single: {1e3 * dt32 / args.steps:.4f} ms/eval")
        batched = None
        if args.replicas > 0:
            reps = synth.hrem_replicas(args.replicas)
            bsteps = max(5, args.steps // 10)
            bm = time_batched(ctx, reps, 1, bsteps, 2)
            log(f"batched mixed x{args.replicas}: {1e3 * bm['dt'] / bsteps:.3f} ms/step, tuning {bm['tuning']}")
            # replica 0 alone in FP64: the force error of the batched mixed path
            rs = resident(ctx, reps[0], 0)
            rs.eval(rebuild_list=True)
            f0 = rs.results()["f"]
            rs.close()
            bd = time_batched(ctx, reps, 0, bsteps, 2, tune=False)
            log(f"batched fp64 x{args.replicas}: {1e3 * bd['dt'] / bsteps:.3f} ms/step")
            batched = dict(bm=bm, bd=bd, bsteps=bsteps, err=relrms(bm["res0"]["f"], f0),
                           err64=relrms(bd["res0"]["f"], f0))

    evals64 = args.steps / dt64
    evals32 = args.steps / dt32
    f64 = r64["f"]
    err32 = relrms(r32["f"], f64)
    kms = st64["pair_kernel"]
    achieved = FLOP_PER_PAIR * npairs / (kms * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": "k_pairs<double,3>", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
            "per_launch": {"pairs": npairs, "flop_per_pair": FLOP_PER_PAIR, "ms": kms,
                           "note": "FP64 VALU (no dense contraction): peak is the FP64 vector rate"}}

    def rnd(d):
        return {k: round(float(v), 4) for k, v in d.items()}

    out = {
        "metric": "SPME+direct force evals/s (24.5k atoms)",
        "value": evals64,
        "unit": "force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt64 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "solvated solute 24,460 atoms, L=61 A, PME 64^3 order 5, alpha 0.35, rc 10 A "
                               "(+0.3 switch), list rebuilt every evaluation", "natoms": s.natoms,
                   "groups": s.ngrp, "evaluated_pairs": npairs, "device": dev},
        "stage_ms": rnd(st64),
        "mixed": {"value": evals32, "unit": "force evals/s", "dtype": "f32 pair/spline/FFT + f64 tables/reductions",
                  "force_relrms_vs_f64": err32, "ms_per_step": 1e3 * dt32 / args.steps, "stage_ms": rnd(st32)},
        "roofline": roof,
    }
    if batched:
        R, K = args.replicas, batched["bsteps"]
        bm, bd = batched["bm"], batched["bd"]
        agg = R * K / bm["dt"]
        out["batched"] = {
            "workload": f"{R} HREM replicas of the 24,460-atom system (solute q*sqrt(s), eps*s, s 1->0.3), "
                        "one launch per stage for all replicas, HIP graph replay, list rebuilt every evaluation",
            "replicas": R, "steps": K, "mixed_evals_per_s": agg, "mixed_ms_per_step": 1e3 * bm["dt"] / K,
            "fp64_evals_per_s": R * K / bd["dt"], "fp64_ms_per_step": 1e3 * bd["dt"] / K,
            "speedup_mixed_batched_vs_plain_fp64": agg / evals64,
            "force_relrms_mixed_vs_f64": batched["err"], "force_relrms_batched_f64_vs_alone": batched["err64"],
            "tuning": bm["tuning"], "stage_ms_mixed": rnd(bm["stages"]), "device_bytes": bm["device_bytes"]}
    if not args.no_cpu:
        nthr = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, nsample, cref = cpu_baseline(s, corr, nthr)
        out["cpu_baseline"] = {"value": v, "unit": "force evals/s", "cores": nthr, "kind": "port",
                               "sample": f"{nsample} full FP64 evaluations (list + force + fft_pme + ferrf) of the "
                                         f"same 24,460-atom system, OpenMP C oracle",
                               "gpu_vs_cpu_force_relrms": relrms(f64, cref["f"])}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
