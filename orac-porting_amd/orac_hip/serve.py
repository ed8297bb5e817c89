"""Continuous-batching server of MD segment requests on one GPU.

A request is a system (a solute in a water box) and a number of outer
r-RESPA steps.  The server keeps every admitted request's coordinates,
velocities, lists and grids resident in HBM and advances all of them
together: each tick is ``chunk`` outer steps of one batch that holds every
active request, one kernel launch per stage for all of them.  Between ticks
finished requests leave (their final state is copied out) and waiting
requests are admitted, in arrival order, while the device bytes of the
active set stay within the HBM budget (each request is charged the larger of
its estimated and its actual device bytes).  A request's trajectory does not
depend on what it is batched with: each system keeps its own step counter
and list schedule, and forces of already running systems are never
recomputed when the batch changes.

``serving_request`` builds the request systems of the serving workload:
a lattice solute of n atoms (jittered, chain-connected) in a box of rigid
waters, grid spacing at most 1 A rounded up to 2^a 3^b 5^c, order 5.
"""
from __future__ import annotations

import dataclasses
import math
import os
import time

import numpy as np

from . import formats
from . import md as _md
from . import synth
from .api import Batch, ResidentSystem


# ---------------------------------------------------------------- workload

def fft_size(n):
    """Smallest 2^a 3^b 5^c >= n."""
    m = max(8, int(math.ceil(n)))
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def _lattice_solute(rng, n, a=2.5, jitter=0.12):
    """n points of a cubic lattice (spacing a) nearest the origin, jittered,
    in snake order so that chain neighbours are lattice neighbours."""
    r = (3.0 * n * a ** 3 / (4.0 * math.pi)) ** (1.0 / 3.0) + 2.0 * a
    m = int(math.ceil(r / a)) + 1
    g = np.arange(-m, m + 1)
    ix, iy, iz = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    d2 = (pts.astype(float) ** 2).sum(axis=1)
    keep = np.argsort(d2, kind="stable")[:n]
    p = pts[keep]
    ys = np.where(p[:, 0] % 2 == 0, p[:, 1], -p[:, 1])
    zs = np.where((p[:, 0] + p[:, 1]) % 2 == 0, p[:, 2], -p[:, 2])
    p = p[np.lexsort((zs, ys, p[:, 0]))].astype(float) * a
    p = p + rng.uniform(-jitter, jitter, size=p.shape)
    radius = math.sqrt(float(d2[keep].max())) * a + jitter
    return p, radius


def serving_request(seed, n_solute, rcut=9.5, alpha=0.35, order=5, shell=11.0, max_bond=3.2, ntypes=64,
                    sigma=(2.0, 2.6)):
    """A solute of n_solute lattice atoms in a cubic box of rigid waters.

    The solute sits on a jittered cubic lattice of 2.5 A (0.064 atoms/A^3, a
    protein's heavy-atom density), so that with LJ sigma U[2.0, 2.6] A it
    relaxes in a few hundred steepest-descent steps.  Charges ~N(0, 0.35) e
    (net 0); ntypes LJ types (sigma as above, eps U[0.02, 0.2] kcal/mol), one
    drawn per atom, so the type tables stay small for any solute size; chain
    1-2 and 1-3 exclusions along the snake (links longer than max_bond A are
    dropped).  The box edge leaves `shell` A of water around the solute
    sphere; waters on a jittered lattice outside it.
    """
    rng = np.random.default_rng(seed)
    pos, radius = _lattice_solute(rng, n_solute)
    q = rng.normal(0.0, 0.35, n_solute)
    q -= q.mean()
    nt = min(ntypes, n_solute)
    sig = rng.uniform(sigma[0], sigma[1], nt)
    eps = rng.uniform(0.02, 0.2, nt)
    types = rng.integers(0, nt, n_solute)
    link = np.sqrt(((pos[1:] - pos[:-1]) ** 2).sum(axis=1)) < max_bond
    i = np.arange(n_solute - 1)[link]
    bonds = np.stack([i, i + 1])
    j = np.arange(n_solute - 2)[link[:-1] & link[1:]]
    bends = np.stack([j, j + 2])
    L = 2.0 * (radius + shell)
    a = 3.1067
    nside = int(L // a)
    a = L / nside
    g = (np.arange(nside) + 0.5) * a - L / 2.0
    o = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    o = o[np.sqrt((o * o).sum(axis=1)) > radius + 1.6]
    o = o + rng.uniform(-0.3, 0.3, size=o.shape)
    sites = synth._place_waters(rng, o)
    K = fft_size(L / 1.0)
    return synth._assemble(L, pos, q, sig, eps, sites, alpha=alpha, nfft=(K, K, K), order=order, rcut=rcut,
                           solute_bonds=bonds, solute_bends=bends, name=f"serve{n_solute}", solute_types=types)


def poisson_arrivals(rng, n, rate):
    """Arrival times (s) of n requests of a Poisson stream of the given rate."""
    return np.cumsum(rng.exponential(1.0 / rate, n))


# ---------------------------------------------------------------- server

@dataclasses.dataclass
class Request:
    rid: int
    system: synth.System
    nsteps: int
    arrival: float = 0.0
    temperature: float = 300.0
    seed: int = 0
    positions: np.ndarray = None    # Cartesian start positions (3, N); default: the system's own
    velocities: np.ndarray = None   # start velocities (3, N); default: Maxwell-Boltzmann(temperature, seed)
    # filled in by the server
    admitted: float = None
    finished: float = None
    final_positions: np.ndarray = None
    final_velocities: np.ndarray = None
    energies: dict = None
    device_bytes: int = 0

    @property
    def latency(self):
        return self.finished - self.arrival


class _Active:
    def __init__(self, ctx, req, precision):
        s = req.system
        cp = np.concatenate([s.bonds, s.bends], axis=1) if s.bonds is not None else np.zeros((2, 0), np.int32)
        self.req = req
        self.rs = ResidentSystem(ctx, s, precision=precision, corr_pairs=cp, corr_fudge=np.ones(cp.shape[1]))
        self.dyn = _md.Dynamics(self.rs, s)
        if req.positions is not None:
            self.dyn.set_positions(req.positions)
        if req.velocities is not None:
            self.dyn.set_velocities(req.velocities)
        else:
            self.dyn.maxwell_boltzmann(req.temperature, req.seed)
        self.left = req.nsteps
        self.e0, self.s0 = None, 0
        req.device_bytes = self.rs.device_bytes()

    def close(self):
        self.rs.close()


def estimate_bytes(s):
    """Device bytes a request will hold (the neighbour list dominates)."""
    K = np.prod(s.nfft)
    Kc = (s.nfft[0] // 2 + 1) * s.nfft[1] * s.nfft[2]
    rl = s.rcut + s.rtol + s.rnei
    nb = 4.0 / 3.0 * math.pi * rl ** 3 * s.ngrp / s.L ** 3
    lst = 4 * s.ngrp * (1.6 * nb + 64.0)
    return int(lst + 8 * K * 2 + 16 * Kc + 400 * s.natoms + (1 << 20))


class Server:
    """Continuous batching of MD segments on one device (see the module doc)."""

    def __init__(self, ctx, hbm_budget, max_active=64, chunk=10, dt=2.0, ninner=4, list_every=5, precision=1,
                 autotune=False, clock=time.perf_counter, restart_dir=None, tuning=None, min_batch_s=0.0):
        self.ctx, self.budget, self.max_active, self.chunk = ctx, int(hbm_budget), max_active, chunk
        self.dt, self.ninner, self.list_every, self.precision = dt, ninner, list_every, precision
        self.autotune, self.clock = autotune, clock
        self.tuning = dict(tuning) if tuning else None  # launch knobs every batch takes (Batch.set_tuning)
        # a running batch takes newcomers only after it has run this long
        # (each new batch primes its newcomers' forces and captures a new step
        # graph; a burst of arrivals then joins in few batches, not one each)
        self.min_batch_s = float(min_batch_s)
        self.restart_dir = restart_dir  # finished requests are checkpointed there in ORAC's restart format
        self.log = []
        self.progress = None  # optional callable(str): admission and completion events

    def run(self, requests):
        """Serve the requests (arrival times relative to the start); returns stats."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        queue, active, used = [], [], 0
        batch = None
        ticks = steps_done = 0
        # seconds forming batches, integrating (the first tick of a new batch
        # apart: it also primes the forces of new systems and captures the
        # step graph), and the rest (admission, results)
        t_batch = t_steps = t_first = 0.0
        fresh = False
        formed = 0.0
        t0 = self.clock()
        while pending or queue or active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            changed = False
            hold = batch is not None and self.clock() - t0 - formed < self.min_batch_s
            while queue and len(active) < self.max_active and not hold:
                need = estimate_bytes(queue[0].system)
                if active and used + need > self.budget:
                    break  # the head waits for memory (arrival order is kept)
                req = queue.pop(0)
                req.admitted = self.clock() - t0
                a = _Active(self.ctx, req, self.precision)
                a.charge = max(need, req.device_bytes)  # the larger of the estimate and the bytes held
                used += a.charge
                active.append(a)
                if self.progress:
                    self.progress(f"t={req.admitted:.3f}s admit request {req.rid} ({req.system.natoms} atoms), "
                                  f"{len(active)} active, {used / 2**30:.2f} GiB charged")
                changed = True
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (self.clock() - t0), 0.05)))
                continue
            if changed or batch is None:
                tb = self.clock()
                if batch is not None:
                    batch.close()
                batch = Batch(self.ctx, [a.rs for a in active])
                if self.tuning:
                    batch.set_tuning(**self.tuning)
                elif self.autotune:
                    batch.autotune(reps=2, rebuild_list=False)
                t_batch += self.clock() - tb
                fresh = True
                formed = self.clock() - t0
            n = min(self.chunk, min(a.left for a in active))
            ts = self.clock()
            _md.run_batch(batch, n, dt=self.dt, ninner=self.ninner, list_every=self.list_every, graph=True)
            self.ctx.sync()
            if fresh:
                t_first += self.clock() - ts
                fresh = False
            else:
                t_steps += self.clock() - ts
            ticks += 1
            steps_done += n * len(active)
            self.log.append((self.clock() - t0, len(active), used))
            done = [a for a in active if a.left == n]
            for a in active:
                a.left -= n
                if a.e0 is None and a.left > 0:
                    # NVE reference point: the total energy after the request's first chunk
                    a.e0, a.s0 = a.dyn.energies()["total"], a.req.nsteps - a.left
            if done:
                for a in done:
                    r = a.req
                    r.final_positions, r.final_velocities = a.dyn.state()
                    r.energies = a.dyn.energies()
                    r.finished = self.clock() - t0
                    if a.e0 is not None and r.nsteps > a.s0:
                        ns = (r.nsteps - a.s0) * self.dt * 1e-6
                        r.energies["drift_kj_mol_per_ns_per_atom"] = float(
                            (r.energies["total"] - a.e0) * synth.EFACT / 1000.0 / ns / r.system.natoms)
                    if self.restart_dir:
                        path = os.path.join(self.restart_dir, f"request{r.rid:05d}.rst")
                        formats.write_restart(path, formats.restart_state(r.system, r.final_positions,
                                                                          r.final_velocities, r.nsteps,
                                                                          r.energies["temperature"]))
                    if self.progress:
                        self.progress(f"t={r.finished:.3f}s done request {r.rid}: latency {r.latency * 1e3:.0f} ms, "
                                      f"T {r.energies['temperature']:.0f} K")
                active = [a for a in active if a.left > 0]
                batch.close()
                batch = None
                for a in done:
                    used -= a.charge
                    a.close()
        wall = self.clock() - t0
        lat = np.array([r.latency for r in requests])
        sim_ns = sum(r.nsteps for r in requests) * self.dt * 1e-6
        return dict(requests=len(requests), wall_s=wall, ticks=ticks, system_steps=steps_done,
                    p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                    mean_ms=float(lat.mean() * 1e3), aggregate_ns_per_day=sim_ns / wall * 86400.0,
                    max_active=max((k for _, k, _ in self.log), default=0),
                    peak_bytes=max((u for _, _, u in self.log), default=0),
                    wall_split_s={"forming_batches": t_batch, "first_tick_of_a_batch": t_first,
                                  "integrating": t_steps, "other": wall - t_batch - t_first - t_steps})


def relax_requests(ctx, requests, niter=300, dmax=0.03, precision=1, equilibrate=0, group_bytes=48 << 30,
                   list_every=5):
    """Start states for each request (preparation, not served time): the
    generators place atoms without regard to LJ overlap, so niter
    steepest-descent steps first; then, with equilibrate > 0, that many NVE
    steps from Maxwell-Boltzmann velocities at the request's temperature with
    the velocities rescaled to it every 10 steps, and the request starts from
    the equilibrated positions and velocities.  Requests are prepared in
    batches of at most group_bytes estimated device bytes; a request's start
    state does not depend on its batch-mates."""
    groups, cur, used = [], [], 0
    for r in requests:
        need = estimate_bytes(r.system)
        if cur and used + need > group_bytes:
            groups.append(cur)
            cur, used = [], 0
        cur.append(r)
        used += need
    if cur:
        groups.append(cur)
    for g in groups:
        act = [_Active(ctx, dataclasses.replace(r, positions=None, velocities=None), precision) for r in g]
        b = Batch(ctx, [a.rs for a in act])
        _md.relax_batch(b, niter, dmax=dmax)
        if equilibrate > 0:
            for a, r in zip(act, g):
                a.dyn.maxwell_boltzmann(r.temperature, r.seed)
            temps = {r.temperature for r in g}
            if len(temps) != 1:
                raise ValueError("requests prepared together share one temperature")
            _md.equilibrate(b, [a.dyn for a in act], temps.pop(), equilibrate, list_every=list_every)
        for a, r in zip(act, g):
            if equilibrate > 0:
                r.positions, r.velocities = a.dyn.state()
            else:
                r.positions, _ = a.dyn.state()
        b.close()
        for a in act:
            a.close()
