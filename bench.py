"""Throughput of the ORAC nonbonded force evaluation on one MI355X.

Workload: the 24,460-atom solvated-solute system (61 A box, 1,660-atom
solute with chain exclusions + 7,600 rigid waters; alpha 0.35, PME grid 64^3,
order 5, cutoff 10 A switched over 0.3 A), synthetic coordinates from the
seeded generator.  One step = one full force evaluation with all state in
HBM: group and molecule centres, linked-cell list rebuild, direct-space
Ewald-erfc + LJ pairs, SPME (spread, r2c FFT, influence function, c2r FFT,
gather) and the per-step exclusion corrections.

Prints one JSON line.  ``value`` is FP64 force evaluations per second; the
mixed-precision path is reported beside it, with its force error against
the FP64 result.  ``roofline`` describes the dominant kernel with stage
durations measured by HIP events on the context stream.  ``cpu_baseline``
times the FP64 CPU oracle (OpenMP C restatement of the reference routines)
on this host on a bounded sample of the same workload.
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


def time_gpu(ctx, s, precision, steps, warmup, corr):
    from orac_hip.api import ResidentSystem
    rs = ResidentSystem(ctx, s, precision=precision, corr_pairs=corr[0], corr_fudge=corr[1])
    for _ in range(warmup):
        rs.eval(rebuild_list=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        rs.eval(rebuild_list=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    res = rs.results()
    # stage durations with HIP events on the context stream (separate loop)
    rs.profile(True)
    stages = np.zeros(5)
    nprof = max(5, min(steps, 20))
    for _ in range(nprof):
        rs.eval(rebuild_list=True)
        stages += np.array(rs.stage_ms())
    stages /= nprof
    rs.profile(False)
    rs.close()
    return dt, res, stages


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (one system or one batch per device); --gpus {args.gpus} is not "
              "supported: sharding across GPUs is not part of this build.", file=sys.stderr)
        sys.exit(2)

    from orac_hip import synth
    from orac_hip.api import Context

    s = synth.config("solvated24k")
    corr = (np.concatenate([s.bonds, s.bends], axis=1), np.ones(s.bonds.shape[1] + s.bends.shape[1]))
    npairs = evaluated_pairs(s)

    with Context(0) as ctx:
        dev = ctx.device_name()
        dt64, r64, st64 = time_gpu(ctx, s, 0, args.steps, args.warmup, corr)
        dt32, r32, st32 = time_gpu(ctx, s, 1, args.steps, args.warmup, corr)

    evals64 = args.steps / dt64
    evals32 = args.steps / dt32
    f64 = r64["f"]
    err32 = float(np.sqrt(((r32["f"] - f64) ** 2).sum() / (f64 ** 2).sum()))
    stage_names = ["list", "centres", "pairs", "pme", "corrections"]
    # dominant stage of the FP64 path
    dom = int(np.argmax(st64))
    if stage_names[dom] == "pairs":
        t = st64[2] * 1e-3
        achieved = FLOP_PER_PAIR * npairs / t / 1e12
        roof = {"bound": "mfma", "kernel": "k_pairs<double>", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "per_launch": {"pairs": npairs, "flop_per_pair": FLOP_PER_PAIR, "ms": st64[2]}}
    else:
        N, G = s.natoms, int(np.prod(s.nfft))
        Gc = (s.nfft[0] // 2 + 1) * s.nfft[1] * s.nfft[2]
        nbytes = (32 * N + 8 * G) + (2 * (8 * G + 16 * Gc) + 32 * Gc) + (8 * G + 56 * N)
        t = st64[3] * 1e-3
        achieved = nbytes / t / 1e9
        roof = {"bound": "hbm", "kernel": "pme stages", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "per_launch": {"bytes": nbytes, "ms": st64[3]}}

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
        "stage_ms": dict(zip(stage_names, [round(float(x), 4) for x in st64])),
        "mixed": {"value": evals32, "unit": "force evals/s", "dtype": "f32 pair/spline/FFT + f64 tables/reductions",
                  "force_relrms_vs_f64": err32, "ms_per_step": 1e3 * dt32 / args.steps,
                  "stage_ms": dict(zip(stage_names, [round(float(x), 4) for x in st32]))},
        "roofline": roof,
    }
    if not args.no_cpu:
        nthr = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, nsample, cref = cpu_baseline(s, corr, nthr)
        err = float(np.sqrt(((f64 - cref["f"]) ** 2).sum() / (cref["f"] ** 2).sum()))
        out["cpu_baseline"] = {"value": v, "unit": "force evals/s", "cores": nthr, "kind": "port",
                               "sample": f"{nsample} full FP64 evaluations (list + force + fft_pme + ferrf) of the "
                                         f"same 24,460-atom system, OpenMP C oracle",
                               "gpu_vs_cpu_force_relrms": err}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
