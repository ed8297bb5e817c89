"""Throughput of the ORAC nonbonded force evaluation on one MI355X.

Workload: the 24,460-atom solvated-solute system (61 A box, 1,660-atom
solute with chain exclusions + 7,600 rigid waters; alpha 0.35, PME grid 64^3,
order 5, cutoff 10 A switched over 0.3 A), synthetic coordinates from the
seeded generator.  One step = one full force evaluation with all state in
HBM: group and molecule centres, linked-cell list rebuild, direct-space
Ewald-erfc + LJ pairs, SPME (spread, r2c FFT, influence function, c2r FFT,
gather) and the per-step exclusion corrections.

Prints one JSON line.  ``value`` is FP64 force evaluations per second of
the single system (the plain path).  ``dynamics`` gives ns/day of on-device
NVE r-RESPA runs (plain FP64 single system; 32 replicas batched in mixed
precision) with their NVE energy drift; ``serving`` the p50/p99 latency and
aggregate ns/day of a continuous-batching run of MD segment requests.  Beside it: the same system in mixed
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
# HBM bytes per launch of k_pairs<double,3,false> on the 24,460-atom system, from the
# committed FETCH_SIZE / WRITE_SIZE profile (rocprof_summary traffic)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "bytes_k_pairs_fp64_solvated24k.txt")
PAIR_KERNEL = "k_pairs<double,3,false>"


def pme_stage_roofline(s, stages, fp64):
    """HBM fraction of the reciprocal-sum stages of one evaluation: algorithmic
    bytes (spread 32N + 8G in FP64, 16N + 8G mixed; FFT stage 2(8G + 16Gc) +
    32Gc; gather 8G + 56N in FP64, 8G + 40N mixed) over the stage's event time."""
    n = s.natoms
    k1, k2, k3 = s.nfft
    g = k1 * k2 * k3
    gc = (k1 // 2 + 1) * k2 * k3
    b = {"spread": (32 if fp64 else 16) * n + 8 * g, "fft": 2 * (8 * g + 16 * gc) + 32 * gc,
         "gather": 8 * g + (56 if fp64 else 40) * n}
    out = {}
    for k, key in (("spread", "pme_spread"), ("fft", "pme_fft"), ("gather", "pme_gather")):
        ms = float(stages.get(key, 0.0))
        if ms <= 0:
            continue
        gbs = b[k] / (ms * 1e-3) / 1e9
        out[k] = {"bound": "hbm", "bytes": b[k], "ms": ms, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": gbs / HBM_PEAK_GBS}
    return out


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
    """Per-step reciprocal corrections of the solute's flexible 1-2 and 1-3 pairs."""
    if s.bonds is None:
        return np.zeros((2, 0), np.int32), np.zeros(0)
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


GRAPH = True  # replay batched and dynamics steps from HIP graphs (--no-graph: eager launches)


def time_batched(ctx, reps, precision, steps, warmup, tune=True):
    from orac_hip.api import Batch
    systems = [resident(ctx, s, precision) for s in reps]
    b = Batch(ctx, systems)
    # the autotuner times graph replays, so eager runs keep the default tuning
    tuning, tune_ms = (b.autotune(reps=3) if tune and GRAPH else (b.tuning(), None))
    trace = tune_summary(b.tune_log()) if tune_ms is not None else None
    for _ in range(warmup):
        b.eval(rebuild_list=True, graph=GRAPH)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.eval(rebuild_list=True, graph=GRAPH)
    ctx.sync()
    dt = time.perf_counter() - t0
    res0 = systems[0].results()
    stages = profile_stages(b, 3)
    mem = sum(x.device_bytes() for x in systems)
    b.close()
    for x in systems:
        x.close()
    return dict(dt=dt, res0=res0, stages=stages, tuning=tuning, tune_ms=tune_ms, device_bytes=mem, trace=trace)


def tune_summary(log):
    """Per tuning knob and value, the fastest candidate's ms per step among
    those the autotuner timed with that value (e.g. persistent 0 vs 1)."""
    out = {}
    for t, ms in log:
        for k, v in t.items():
            d = out.setdefault(k, {})
            d[str(v)] = min(ms, d.get(str(v), float("inf")))
    return {k: v for k, v in out.items() if len(v) > 1}


def prepare_md(ctx, systems, precision, relax=300, equil=200, temperature=300.0, seed=2024):
    """Steepest descent, Maxwell-Boltzmann velocities, then velocity rescaling to
    the temperature (every 10 steps) while the strained start relaxes."""
    from orac_hip import md
    from orac_hip.api import Batch
    rss = [resident(ctx, s, precision) for s in systems]
    dyns = [md.Dynamics(rs, s) for rs, s in zip(rss, systems)]
    b = Batch(ctx, rss)
    md.relax_batch(b, relax, dmax=0.03)
    for k, d in enumerate(dyns):
        d.maxwell_boltzmann(temperature, seed + k)
    md.equilibrate(b, dyns, temperature, equil, graph=GRAPH)
    return rss, dyns, b


def time_md(ctx, b, nsteps, warmup, dt=2.0, ninner=4, list_every=5, nmid=None):
    from orac_hip import md
    md.run_batch(b, warmup, dt=dt, ninner=ninner, list_every=list_every, graph=GRAPH, nmid=nmid)
    ctx.sync()
    t0 = time.perf_counter()
    md.run_batch(b, nsteps, dt=dt, ninner=ninner, list_every=list_every, graph=GRAPH, nmid=nmid)
    ctx.sync()
    return time.perf_counter() - t0


def time_md2(ctx, b, nsteps, warmup):
    """time_md for systems with two shells: outer step 4 fs, inner shell twice
    per step, fast terms four times per inner step, lists every 3 outer steps."""
    return time_md(ctx, b, nsteps, warmup, dt=4.0, ninner=4, list_every=3, nmid=2)


def nve_drift2(ctx, b, dyn, nsteps):
    return nve_drift(ctx, b, dyn, nsteps, dt=4.0, ninner=4, list_every=3, nmid=2)


def nve_drift(ctx, b, dyn, nsteps, dt=2.0, ninner=4, list_every=5, nmid=None):
    """Total-energy drift (kJ/mol per ns) of the first system over nsteps: the
    slope of a linear fit to the total energy sampled every 10 steps."""
    from orac_hip import md, synth
    e0 = dyn.energies()
    tot, t_ns = [e0["total"]], [0.0]
    for k in range(max(1, nsteps // 10)):
        md.run_batch(b, 10, dt=dt, ninner=ninner, list_every=list_every, graph=GRAPH, nmid=nmid)
        tot.append(dyn.energies()["total"])
        t_ns.append((k + 1) * 10 * dt * 1e-6)
    ctx.sync()
    e1 = dyn.energies()
    kj = synth.EFACT / 1000.0
    slope = float(np.polyfit(np.array(t_ns), np.array(tot) * kj, 1)[0])
    return dict(kj_mol_per_ns=slope,
                kj_mol_per_ns_per_atom=slope / dyn.s.natoms, energy_sd_kj_mol=float(np.std(np.array(tot) * kj)),
                rel_per_ns=slope / (abs(e0["total"]) * kj),
                temperature=(e0["temperature"], e1["temperature"]), steps=nsteps, fail=e1["fail"])


def ns_per_day(steps, wall, dt=2.0):
    return steps * dt * 1e-6 / wall * 86400.0


def _served_physics(reqs):
    """Per-request final temperature and NVE drift (kJ/mol/ns per atom) of a served run."""
    temps = np.array([r.energies["temperature"] for r in reqs])
    drift = np.array([r.energies.get("drift_kj_mol_per_ns_per_atom", np.nan) for r in reqs])
    return dict(final_temperature_K={"min": float(temps.min()), "median": float(np.median(temps)),
                                     "max": float(temps.max())},
                nve_drift_kj_mol_per_ns_per_atom={"median_abs": float(np.nanmedian(np.abs(drift))),
                                                  "max_abs": float(np.nanmax(np.abs(drift)))})


def serving_leg(ctx, nreq=64, rate=200.0, nsteps=500, budget_gib=16, seed=4242, equil=600, tuning=None):
    """Poisson stream of MD segment requests (solutes log-uniform in 2k-150k
    atoms), each relaxed and equilibrated at 300 K before it is submitted;
    the HBM budget is small enough that requests queue for memory."""
    from orac_hip import serve
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(2000), np.log(150000), nreq)).astype(int)
    arrivals = serve.poisson_arrivals(rng, nreq, rate)
    reqs = [serve.Request(rid=k, system=serve.serving_request(seed + k, int(n), rcut=float(rng.uniform(9.0, 10.0))),
                          nsteps=nsteps, arrival=float(arrivals[k]), seed=seed + k) for k, n in enumerate(sizes)]
    log(f"serving: {nreq} requests, {sum(r.system.natoms for r in reqs)} atoms in all; relaxing and equilibrating")
    serve.relax_requests(ctx, reqs, niter=200, equilibrate=equil)
    srv = serve.Server(ctx, hbm_budget=budget_gib << 30, chunk=25, tuning=tuning, min_batch_s=0.1)
    srv.progress = log
    st = srv.run(reqs)
    waits = np.array([r.admitted - r.arrival for r in reqs])
    st.update(sizes_atoms=[r.system.natoms for r in reqs], rate_per_s=rate, steps_per_request=nsteps,
              budget_gib=budget_gib, fail=sum(r.energies["fail"] for r in reqs),
              queued_requests=int((waits > 0.05).sum()), max_admission_wait_ms=float(waits.max() * 1e3),
              **_served_physics(reqs))
    return st


def served_hrem_leg(ctx, reps, nsteps=500, rate=200.0, seed=77, equil=400, tuning=None):
    """The replicas of the batched leg as served requests: Poisson arrivals,
    continuous batching in mixed precision, each request equilibrated at 300 K
    beforehand; aggregate ns/day and p50/p99 of the segment latency."""
    from orac_hip import serve
    rng = np.random.default_rng(seed)
    arrivals = serve.poisson_arrivals(rng, len(reps), rate)
    reqs = [serve.Request(rid=k, system=s, nsteps=nsteps, arrival=float(arrivals[k]), seed=seed + k)
            for k, s in enumerate(reps)]
    serve.relax_requests(ctx, reqs, niter=300, equilibrate=equil)
    srv = serve.Server(ctx, hbm_budget=64 << 30, chunk=50, tuning=tuning, min_batch_s=0.1)
    st = srv.run(reqs)
    st.update(replicas=len(reps), steps_per_request=nsteps, rate_per_s=rate,
              fail=sum(r.energies["fail"] for r in reqs), **_served_physics(reqs))
    return st


def c0_leg(ctx, steps, md=True):
    """The 648-atom water box: one FP64 evaluation against the mixed one, and
    (md) 100 outer r-RESPA steps in FP64 (ns/day, NVE drift)."""
    from orac_hip import synth
    from orac_hip.api import ResidentSystem
    w = synth.config("water216")
    res = {}
    for prec in (0, 1):
        rs = ResidentSystem(ctx, w, precision=prec)
        rs.eval(rebuild_list=True)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            rs.eval(rebuild_list=True)
        ctx.sync()
        res[prec] = (time.perf_counter() - t0, rs.results())
        rs.close()
    out = {"workload": "216 rigid waters (648 atoms), L=18.64 A, PME 20^3 order 5, alpha 0.40, rc 8.5 A",
           "fp64_evals_per_s": steps / res[0][0], "mixed_evals_per_s": steps / res[1][0],
           "force_relrms_mixed_vs_f64": relrms(res[1][1]["f"], res[0][1]["f"])}
    if not md:
        return out
    rss, dyns, b = prepare_md(ctx, [w], 0, equil=300)
    wall = time_md(ctx, b, 100, 5, list_every=5)
    # drift over 2 ps (1,000 steps of 2 fs; a 0.6 s run): a shorter window
    # measures the total energy's fluctuation, not its drift
    drift = nve_drift(ctx, b, dyns[0], 1000, list_every=5)
    b.close()
    rss[0].close()
    out["md_100_steps"] = {"ns_per_day": ns_per_day(100, wall), "ms_per_step": 10 * wall, "nve": drift}
    return out


DUDL_H = 0.005


def dudl_central(b, rss, lams):
    """dU/dlambda of each window (internal units) from its dwrk = U(lambda) -
    U(lambda0) at two lambda0 around lambda: (dwrk(l - h) - dwrk(l + h)) / 2h,
    and at the ends of [0, 1] the one-sided three-point forms
    (dwrk(2h) - 4 dwrk(h)) / 2h and (4 dwrk(1 - h) - dwrk(1 - 2h)) / 2h.
    The windows' own lambda0 are restored afterwards."""
    from orac_hip import synth
    h = DUDL_H
    plans = []
    for lam in lams:
        l = float(lam[1])
        if l - h < 0.0:
            plans.append(((l + 2 * h, 1.0), (l + h, -4.0)))
        elif l + h > 1.0:
            plans.append(((l - h, 4.0), (l - 2 * h, -1.0)))
        else:
            plans.append(((l - h, 1.0), (l + h, -1.0)))
    du = np.zeros(len(lams))
    for k in range(2):
        for rs, lam, pl in zip(rss, lams, plans):
            rs.set_lambda(synth.lambda_state(float(lam[1]), lrm0=pl[k][0]))
        b.eval(rebuild_list=True)
        for w, (rs, pl) in enumerate(zip(rss, plans)):
            du[w] += pl[k][1] * rs.alchemy_results()["dwrk"] / (2 * h)
    for rs, lam in zip(rss, lams):
        rs.set_lambda(lam)
    return du.tolist()


def alchemy_leg(ctx, steps, md_steps=0):
    """16 lambda-windows of a 48-atom ligand decoupled in a 40 A water box
    (6,399 atoms, PME 48^3 order 5, rc 9 A), every window resident with its
    own charge grids, one batch: soft-core direct pairs, reciprocal sum on the
    window's charges and the energy-only grid of its neighbour window's
    charges (dwrk = U(lambda_w) - U(lambda_w')).  Mixed against FP64: forces
    and per-window dU/dlambda (dudl_central)."""
    from orac_hip import md, synth
    from orac_hip.api import Batch, ResidentSystem
    s, al, lams = synth.alchemy_windows()
    cp, cf = corrections(s)
    res = {}
    for prec in (1, 0):
        rss = []
        for lam in lams:
            rs = ResidentSystem(ctx, s, precision=prec, corr_pairs=cp, corr_fudge=cf)
            rs.set_alchemy(al, lam)
            rss.append(rs)
        b = Batch(ctx, rss)
        tuning = b.autotune(reps=3)[0] if (prec == 1 and GRAPH) else b.tuning()
        for _ in range(2):
            b.eval(rebuild_list=True, graph=GRAPH)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.eval(rebuild_list=True, graph=GRAPH)
        ctx.sync()
        dt = time.perf_counter() - t0
        win = [(rs.results()["f"], rs.alchemy_results()) for rs in rss]
        stages = profile_stages(b, 3)
        r = dict(dt=dt, win=win, stages=stages, tuning=tuning, dudl=dudl_central(b, rss, lams))
        if md_steps and prec == 1:
            dyns = [md.Dynamics(rs, s) for rs in rss]
            md.relax_batch(b, 200, dmax=0.03)
            for k, d in enumerate(dyns):
                d.maxwell_boltzmann(300.0, 77 + k)
            md.equilibrate(b, dyns, 300.0, 100, graph=GRAPH)
            r["md_wall"] = time_md(ctx, b, md_steps, 5)
        b.close()
        for rs in rss:
            rs.close()
        res[prec] = r
    n = len(lams)
    kj = synth.EFACT / 1000.0  # kJ/mol per internal energy unit
    du64 = np.array(res[0]["dudl"]) * kj
    du32 = np.array(res[1]["dudl"]) * kj
    ferr = max(relrms(a[0], b[0]) for a, b in zip(res[1]["win"], res[0]["win"]))
    out = {"workload": "16 lambda-windows (charge and soft-core LJ lambda 0->1) of a 48-atom ligand in a 40 A water box "
                       "(6,399 atoms), PME 48^3 order 5, alpha 0.35, rc 9 A; per window its own charge grid and the "
                       "energy-only grid of the next window's charges; one batch, list rebuilt every evaluation",
           "windows": n, "natoms": s.natoms, "steps": steps,
           "mixed_evals_per_s": n * steps / res[1]["dt"], "fp64_evals_per_s": n * steps / res[0]["dt"],
           "mixed_ms_per_step": 1e3 * res[1]["dt"] / steps, "fp64_ms_per_step": 1e3 * res[0]["dt"] / steps,
           "lambda": [float(lam[1]) for lam in lams], "lambda0": [float(lam[3]) for lam in lams],
           "dU_dlambda_fp64": du64.tolist(), "dU_dlambda_mixed": du32.tolist(),
           "dU_dlambda_unit": "kJ/mol per unit lambda",
           "dU_dlambda_method": f"second-order finite difference of dwrk at lambda0 = lambda +- {DUDL_H} "
                                "(one-sided three-point at lambda 0 and 1), both precisions",
           "dU_dlambda_max_abs_err_mixed": float(np.abs(du32 - du64).max()),
           "dU_dlambda_max_rel_err_mixed": float(np.abs(du32 - du64).max() / np.abs(du64).max()),
           "force_relrms_mixed_vs_f64_max": ferr, "tuning": res[1]["tuning"],
           "stage_ms_mixed": {k: round(float(v), 4) for k, v in res[1]["stages"].items()}}
    if "md_wall" in res[1]:
        out["md_mixed"] = {"steps": md_steps, "ms_per_step": 1e3 * res[1]["md_wall"] / md_steps,
                           "ns_per_day_aggregate": ns_per_day(md_steps * n, res[1]["md_wall"])}
    return out


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


def cpu_reference(s, corr, budget_s=20.0):
    """The reference's own Fortran (oracle/_ref, compiled in place from the
    reference sources, serial, fftpack branch) on the same evaluation: list,
    force, fft_pme, ferrf; as many evaluations as fit in budget_s (at least one)."""
    import oracle as O
    if not O.ref_available():
        return None
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nc = int(s.L // 3.0)
    nsample, f = 0, None
    t0 = time.perf_counter()
    while True:
        st = O.ref_lc_index(nc, nc, nc, rc, s.co)
        nnlpp = O.ref_lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
        d = O.ref_force(s, nnlpp, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0, fa=fa)
        p = O.ref_fft_pme(s)
        c = O.ref_ferrf(s, corr[0], 1, 1.0)
        f = d["f"] + p["f"] + c["f"]
        nsample += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    dt = time.perf_counter() - t0
    return nsample / dt, nsample, f


def cpu_reference_omp(s, corr, nthr, budget_s=8.0):
    """The reference's own OpenMP build (oracle/_ref/libref_omp.so: -D_OMP_
    -fopenmp, ORAC's per-thread lists and contact maps) on nthr threads, on
    the same evaluation: list, force, fft_pme (fftpack), ferrf; as many
    evaluations as fit in budget_s (at least one)."""
    import oracle as O
    if not O.ref_omp_available():
        return None
    f, done, secs = O.ref_omp_evaluation(s, corr[0], nthr, nsample=1000, budget_s=budget_s)
    return done / secs, done, f


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def relrms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--replicas", type=int, default=32, help="HREM replicas of the batched leg (0 skips it)")
    ap.add_argument("--md-steps", type=int, default=200, help="outer steps of each dynamics leg (0 skips them)")
    ap.add_argument("--serve-requests", type=int, default=64, help="requests of the serving leg (0 skips it)")
    ap.add_argument("--serve-steps", type=int, default=500, help="outer steps per served request")
    ap.add_argument("--serve-hrem", type=int, default=1,
                    help="serve the HREM replicas of the batched leg as requests (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true", help="launch batched and dynamics steps eagerly")
    ap.add_argument("--no-water", action="store_true", help="skip the 648-atom water box leg")
    ap.add_argument("--no-alchemy", action="store_true", help="skip the lambda-window leg")
    ap.add_argument("--only-alchemy", action="store_true",
                    help="run the lambda-window leg alone and print its record (for profiling that workload)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device copy of this size first (calibrates HBM byte counters)")
    args = ap.parse_args()
    # stdout carries only the result line: file descriptor 1 goes to stderr for
    # the whole run (the reference Fortran of the CPU leg writes progress lines
    # to unit 6, flushed as late as process exit) and the line is written to
    # the saved descriptor
    sys.stdout.flush()
    result_fd = os.dup(1)
    os.dup2(2, 1)
    global GRAPH
    GRAPH = not args.no_graph
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (one system or one batch per device); --gpus {args.gpus} is not "
              "supported: the library runs one system or one batch per device and does not shard across GPUs.",
              file=sys.stderr)
        sys.exit(2)

    from orac_hip import synth
    from orac_hip.api import Batch, Context

    s = synth.config("solvated24k")
    corr = corrections(s)
    npairs = evaluated_pairs(s)
    log(f"system: {s.natoms} atoms, {s.ngrp} groups, {npairs} evaluated pairs")

    if args.only_alchemy:
        with Context(0) as ctx:
            rec = alchemy_leg(ctx, args.steps, md_steps=args.md_steps)
        sys.stdout.flush()
        os.write(result_fd, (json.dumps(rec) + "\n").encode())
        return

    with Context(0) as ctx:
        dev = ctx.device_name()
        if args.copy_probe_mb:
            ctx.probe_copy(args.copy_probe_mb << 20)
            log(f"copy probe (k_probe_copy): {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written")
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
        c0 = None
        if not args.no_water:
            c0 = c0_leg(ctx, args.steps, md=args.md_steps > 0)
            log(f"water216: {c0['fp64_evals_per_s']:.0f} FP64 evals/s"
                + (f", md {c0['md_100_steps']['ms_per_step']:.3f} ms/step" if "md_100_steps" in c0 else ""))
        alch = None
        if not args.no_alchemy:
            alch = alchemy_leg(ctx, max(5, args.steps // 10), md_steps=args.md_steps)
            log(f"alchemy x{alch['windows']}: mixed {alch['mixed_ms_per_step']:.3f} ms/step, fp64 "
                f"{alch['fp64_ms_per_step']:.3f} ms/step, dU/dl rel err {alch['dU_dlambda_max_rel_err_mixed']:.2e}")
        dyn = None
        if args.md_steps > 0:
            K = args.md_steps
            # plain FP64, one system at a time (what 32 replicas cost sequentially, per replica)
            rss, dyns, b = prepare_md(ctx, [s], 0)
            w64 = time_md(ctx, b, K, 5)
            drift64 = nve_drift(ctx, b, dyns[0], max(K, 1000))  # 2 ps: drift, not fluctuation
            # the same system with ORAC's nonbonded split in two shells: the
            # outer shell, PME and corrections every 4 fs, the inner shell
            # (5.0 A, switched over 0.5 A) every 2 fs, the fast terms every 0.5 fs
            rss[0].set_shells(5.0, 0.5)
            b.close()
            b = Batch(ctx, rss)
            w2 = time_md2(ctx, b, K, 5)
            drift2 = nve_drift2(ctx, b, dyns[0], max(K // 2, 250))
            b.close()
            rss[0].close()
            log(f"md fp64 single: {1e3 * w64 / K:.3f} ms/step, NVE {drift64}; two shells {1e3 * w2 / K:.3f} ms/4 fs")
            dyn = dict(fp64=dict(ns_per_day=ns_per_day(K, w64), ms_per_step=1e3 * w64 / K, nve=drift64),
                       fp64_two_shells=dict(ns_per_day=ns_per_day(K, w2, dt=4.0), ms_per_outer_step=1e3 * w2 / K,
                                            nve=drift2,
                                            workload="outer shell (5.0-10 A) + PME + corrections every 4 fs, inner "
                                                     "shell (0-5.0 A, switched over 0.5 A) every 2 fs, fast terms "
                                                     "every 0.5 fs, lists every 3 outer steps"))
            if args.replicas > 0:
                rss, dyns, b = prepare_md(ctx, reps, 1)
                if GRAPH:  # the autotuner times graph replays
                    b.autotune(reps=2, rebuild_list=False)
                wm = time_md(ctx, b, K, 5)
                driftm = nve_drift(ctx, b, dyns[0], max(K, 500))
                log(f"md mixed x{args.replicas}: {1e3 * wm / K:.3f} ms/step, NVE {driftm}")
                dyn["batched_mixed"] = dict(replicas=args.replicas, ns_per_day_aggregate=ns_per_day(K * args.replicas,
                                                                                                   wm),
                                            ms_per_step=1e3 * wm / K, nve=driftm, tuning=b.tuning())
                dyn["speedup_batched_mixed_vs_plain_fp64"] = (dyn["batched_mixed"]["ns_per_day_aggregate"]
                                                              / dyn["fp64"]["ns_per_day"])
                b.close()
                for x in rss:
                    x.close()
        served = None
        served_hrem = None
        if args.serve_hrem and args.replicas > 0 and args.md_steps > 0:
            tun = dyn["batched_mixed"]["tuning"] if dyn and "batched_mixed" in dyn else None
            served_hrem = served_hrem_leg(ctx, reps, nsteps=args.serve_steps, tuning=tun)
            log(f"served HREM x{args.replicas}: p50 {served_hrem['p50_ms']:.1f} ms p99 {served_hrem['p99_ms']:.1f} ms, "
                f"{served_hrem['aggregate_ns_per_day']:.1f} ns/day")
            if dyn:
                served_hrem["speedup_vs_plain_fp64_ns_per_day"] = (served_hrem["aggregate_ns_per_day"]
                                                                   / dyn["fp64"]["ns_per_day"])
        if args.serve_requests > 0:
            tun = dyn["batched_mixed"]["tuning"] if dyn and "batched_mixed" in dyn else None
            served = serving_leg(ctx, nreq=args.serve_requests, nsteps=args.serve_steps, tuning=tun)
            log(f"serving: p50 {served['p50_ms']:.1f} ms p99 {served['p99_ms']:.1f} ms, "
                f"{served['aggregate_ns_per_day']:.1f} ns/day")

    evals64 = args.steps / dt64
    evals32 = args.steps / dt32
    f64 = r64["f"]
    err32 = relrms(r32["f"], f64)
    kms = st64["pair_kernel"]
    achieved = FLOP_PER_PAIR * npairs / (kms * 1e-3) / 1e12
    from orac_hip.rocprof_summary import read_traffic, read_stamp
    traffic = read_traffic(TRAFFIC_PROFILE)
    # compute-bound on the FP64 rate: the pair arithmetic is FP64 VALU (no
    # matrix contraction); MI355X's FP64 vector and matrix peaks are the same
    # 78.6 TFLOP/s, the contract's "mfma" compute roof
    roof = {"bound": "mfma", "pipe": "FP64 VALU", "kernel": PAIR_KERNEL, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "per_launch": {"pairs": npairs, "flop_per_pair": FLOP_PER_PAIR, "ms": kms,
                           "traffic_unit": "bytes (FETCH_SIZE + WRITE_SIZE, probe-scaled)",
                           "traffic_source": read_stamp(TRAFFIC_PROFILE),
                           "note": "FP64 VALU (no dense contraction): peak is the FP64 vector rate"},
            "stages": pme_stage_roofline(s, st64, fp64=True)}

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
            "tuning": bm["tuning"], "tuning_candidates_best_ms": bm["trace"],
            "stage_ms_mixed": rnd(bm["stages"]), "device_bytes": bm["device_bytes"]}
    if c0:
        out["water216"] = c0
    if alch:
        out["alchemy"] = alch
    if dyn:
        dyn["workload"] = ("NVE r-RESPA in ORAC's step order: 2 fs outer step, 4 inner steps for the solute harmonic "
                           "terms, rigid waters (ORAC's matrix-inversion SHAKE/RATTLE), list every 5 steps; after 300 "
                           "steepest-descent steps and 200 steps of velocity rescaling at 300 K from "
                           "Maxwell-Boltzmann velocities")
        out["dynamics"] = dyn
    if served_hrem:
        served_hrem["workload"] = (f"the {args.replicas} HREM replicas of the batched leg as served MD segment "
                                   "requests (Poisson arrivals), each relaxed and equilibrated at 300 K, mixed "
                                   "precision, continuous batching")
        out["served_hrem"] = served_hrem
    if served:
        served["workload"] = ("Poisson stream of MD segment requests (lattice solutes of 2k-150k atoms, log-uniform, "
                              "in water boxes; grid spacing <= 1 A; order 5; cutoff 9-10 A), mixed precision, "
                              "continuous batching with HBM-budget admission")
        out["serving"] = served
    if not args.no_cpu:
        # the host's cores, at most 16 (the reference's OpenMP build allows 16 threads, omp_integr mcor3)
        nthr = args.cpu_threads or min(16, os.cpu_count() or 1)
        host = {"cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
                "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
        v, nsample, cref = cpu_baseline(s, corr, nthr)
        port = {"value": v, "unit": "force evals/s", "cores": nthr, "kind": "port",
                "sample": f"{nsample} full FP64 evaluations (list + force + fft_pme + ferrf) of the "
                          f"same 24,460-atom system, OpenMP C oracle",
                "gpu_vs_cpu_force_relrms": relrms(f64, cref["f"])}
        omp = cpu_reference_omp(s, corr, nthr)
        if omp:
            out["cpu_baseline"] = {
                "value": omp[0], "unit": "force evals/s", "cores": nthr, "kind": "reference",
                "sample": f"{omp[1]} full FP64 evaluations (lc_index + lc_list + force + fft_pme + ferrf) of the same "
                          f"24,460-atom system by the reference's own OpenMP build (-D_OMP_ -fopenmp, compiled in "
                          f"place; fftpack PME: FFTW is not in the image) on {nthr} threads",
                "gpu_vs_reference_force_relrms": relrms(f64, omp[2]), "host": host, "port": port}
            if nthr > 8:
                o8 = cpu_reference_omp(s, corr, 8)
                out["cpu_baseline"]["reference_8_threads"] = {"value": o8[0], "unit": "force evals/s", "cores": 8,
                                                              "kind": "reference", "sample": f"{o8[1]} evaluations"}
        else:
            out["cpu_baseline"] = dict(port, host=host)
        ref = cpu_reference(s, corr)
        if ref:
            out["cpu_baseline"]["reference_serial"] = {
                "value": ref[0], "unit": "force evals/s", "cores": 1, "kind": "reference",
                "sample": f"{ref[1]} evaluations by the reference Fortran compiled in place (serial, fftpack "
                          "branch; FFTW is not in the image)",
                "gpu_vs_reference_force_relrms": relrms(f64, ref[2])}
    sys.stdout.flush()
    os.write(result_fd, (json.dumps(out) + "\n").encode())


if __name__ == "__main__":
    main()
