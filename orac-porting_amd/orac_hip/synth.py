"""Seeded synthetic systems in ORAC's internal data contract.

Every array here is what ORAC's nonbonded routines consume, in ORAC's own
conventions, so the same arrays go unchanged to the HIP library, to the C
oracle and to the compiled reference:

* reduced coordinates ``xp = oc . x`` with ``co`` = half the cell matrix, so
  the box spans [-1, 1) (``change_frame.f``, ``setuns.f:43`` boxl=2);
* charges in internal units ``q / sqrt(unitc)`` (``covchg.f:38``);
* packed Lennard-Jones tables ``ecc12 = 4 eps sigma^12``, ``ecc6 = 4 eps sigma^6``
  in internal energy units, ``type_table(i,j) = j(j-1)/2 + i`` for i <= j
  (``covnb.f:81-106``);
* groups ``grppt(2, ngrp)``, ``atomg``, ``atomp``, ``pmass`` (mass / group mass,
  ``redmss.f:15``), group centres ``xpg = sum pmass * xp`` (``appbou.f``),
  molecule centres of mass ``xpcm``;
* the non-bonded exclusion map ``mapnl`` = per atom ``[m, j1..jm]`` with
  ``j > i`` only (``mappa.f:120-130``).

Index arrays are 1-based, as Fortran passes them.  All float arrays are
float64.  Nothing here calls the GPU.
"""
from __future__ import annotations

import dataclasses
import functools
import math

import numpy as np

# ---- units, restated from setuns.f:34-50 -------------------------------
PI = 4.0 * math.atan(1.0)
AVOGAD = 6.022140857e23
ELECHG = 1.60217662e-19
EPSO = 8.854187817e-12
UNITM = 1.0 / (AVOGAD * 1000.0)
UNITL = 1.0e-10
UNITT = 1.0e-15
UNITE = UNITM * (UNITL / UNITT) ** 2
UNITC = 4.0 * PI * EPSO * UNITL * UNITE / (ELECHG * ELECHG)
EFACT = UNITE * AVOGAD  # J/mol per internal energy unit
# covnb.f:96 multiplies by the default-real literal 4.184 (REAL(4)).
KCAL_F32 = float(np.float32(4.184))
MASSINFTY = 1.0e20  # blkdta.f90:280

WATER_Q = (-0.834, 0.417, 0.417)
WATER_MASS = (15.9994, 1.008, 1.008)
R_OH = 0.9572
ANG_HOH = 104.52
O_SIGMA = 3.1507
O_EPS = 0.1521


def charge_internal(q_e):
    return np.asarray(q_e, dtype=np.float64) / math.sqrt(UNITC)


def lj_internal(eps_kcal, sigma):
    """Return (ecc12, ecc6) for one pair, restating covnb.f:94-99."""
    tmpi = 4.0 * 1000.0 * eps_kcal * KCAL_F32 / (UNITE * AVOGAD)
    s6 = sigma ** 6
    return tmpi * s6 * s6, tmpi * s6


@dataclasses.dataclass
class System:
    """One periodic system, ready for the nonbonded force path."""
    L: float
    xp0: np.ndarray
    yp0: np.ndarray
    zp0: np.ndarray
    charge: np.ndarray
    mass: np.ndarray
    nbtype: np.ndarray          # int32, 1-based
    ntypes: int                 # ``ma``
    type_table: np.ndarray      # int32 (ma, ma), Fortran order, 1-based packed index
    ecc12: np.ndarray
    ecc6: np.ndarray
    grppt: np.ndarray           # int32 (2, ngrp) Fortran order, 1-based
    atomg: np.ndarray           # int32, 1-based group of each atom
    atomp: np.ndarray           # int32, 1-based molecule of each atom
    ss_index: np.ndarray        # int32, 1 = solute, 2 = solvent
    pmass: np.ndarray
    co: np.ndarray              # (3,3) Fortran order
    oc: np.ndarray
    volume: float
    alpha: float
    nfft: tuple
    order: int
    rcut: float
    rtol: float = 0.3
    rnei: float = 0.3
    constraints: np.ndarray = None   # (2, n) 1-based; energy-only correction
    bonds: np.ndarray = None         # (2, n) 1-based; per-step correction
    bends: np.ndarray = None         # (2, n) 1-based 1-3 pairs
    name: str = ""

    @property
    def natoms(self):
        return int(self.xp0.shape[0])

    @property
    def ngrp(self):
        return int(self.grppt.shape[1])

    @property
    def nprot(self):
        return int(self.atomp.max())

    def group_centres(self):
        """xpg per appbou.f:19-33 (same summation order)."""
        ng = self.ngrp
        out = np.zeros((3, ng))
        for d, c in enumerate((self.xp0, self.yp0, self.zp0)):
            acc = np.zeros(ng)
            first = self.grppt[0] - 1
            size = self.grppt[1] - self.grppt[0] + 1
            for k in range(int(size.max())):
                sel = k < size
                idx = first[sel] + k
                acc[sel] = acc[sel] + self.pmass[idx] * c[idx]
            out[d] = acc
        return out

    def molecule_coms(self):
        """Molecule centres of mass in reduced coordinates."""
        npr = self.nprot
        out = np.zeros((3, npr))
        mtot = np.bincount(self.atomp - 1, weights=self.mass, minlength=npr)
        for d, c in enumerate((self.xp0, self.yp0, self.zp0)):
            out[d] = np.bincount(self.atomp - 1, weights=self.mass * c, minlength=npr) / mtot
        return out

    def cartesian(self):
        x = self.co[0, 0] * self.xp0 + self.co[0, 1] * self.yp0 + self.co[0, 2] * self.zp0
        y = self.co[1, 0] * self.xp0 + self.co[1, 1] * self.yp0 + self.co[1, 2] * self.zp0
        z = self.co[2, 0] * self.xp0 + self.co[2, 1] * self.yp0 + self.co[2, 2] * self.zp0
        return x, y, z

    def exclusion_pairs(self):
        parts = [p for p in (self.constraints, self.bonds, self.bends) if p is not None and p.size]
        if not parts:
            return np.zeros((2, 0), dtype=np.int32)
        return np.concatenate(parts, axis=1).astype(np.int32)

    def mapnl(self):
        """Exclusion map restating mappa.f: per atom i, [m, j1..jm], j > i ascending."""
        n = self.natoms
        pairs = self.exclusion_pairs()
        lo = np.minimum(pairs[0], pairs[1])
        hi = np.maximum(pairs[0], pairs[1])
        keep = lo != hi
        lo, hi = lo[keep], hi[keep]
        key = np.unique(lo.astype(np.int64) * (n + 1) + hi)
        lo = (key // (n + 1)).astype(np.int64)
        hi = (key % (n + 1)).astype(np.int32)
        cnt = np.bincount(lo - 1, minlength=n)
        out = np.empty(n + len(hi), dtype=np.int32)
        starts = np.concatenate(([0], np.cumsum(cnt + 1)[:-1]))
        out[starts] = cnt
        pos = np.repeat(starts + 1, cnt) + (np.arange(len(hi)) - np.repeat(np.cumsum(cnt) - cnt, cnt))
        out[pos] = hi
        return out

    def with_coords(self, xp0, yp0, zp0):
        return dataclasses.replace(self, xp0=np.ascontiguousarray(xp0), yp0=np.ascontiguousarray(yp0),
                                   zp0=np.ascontiguousarray(zp0))


# ---- geometry helpers ----------------------------------------------------

def _random_rotations(rng, n):
    """Uniform rotations on SO(3) (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _water_template():
    half = math.radians(ANG_HOH) / 2.0
    h1 = np.array([R_OH * math.sin(half), R_OH * math.cos(half), 0.0])
    h2 = np.array([-R_OH * math.sin(half), R_OH * math.cos(half), 0.0])
    return np.stack([np.zeros(3), h1, h2])


def _place_waters(rng, o_pos):
    """Rigid 3-site waters with O at o_pos (n,3) and uniform random orientation."""
    R = _random_rotations(rng, len(o_pos))
    t = _water_template()
    sites = np.einsum("nij,aj->nai", R, t) + o_pos[:, None, :]
    return sites.reshape(-1, 3)


def _rsa_sphere(rng, n, radius, dmin):
    """Random sequential addition of n points in a sphere, pairwise >= dmin."""
    cell = dmin
    ncell = int(math.ceil(2 * radius / cell)) + 2
    grid = {}
    pts = np.empty((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > 200 * n:
            raise RuntimeError("sphere packing failed")
        p = (rng.random(3) * 2 - 1) * radius
        if p @ p > radius * radius:
            continue
        c = tuple(((p + radius) // cell).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((c[0] + dx, c[1] + dy, c[2] + dz), ()):
                        d = pts[j] - p
                        if d @ d < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        pts[k] = p
        grid.setdefault(c, []).append(k)
        k += 1
    del ncell
    return pts


def _chain_exclusions(pos):
    """Greedy nearest-unvisited chain through pos: 1-2 (bonds) and 1-3 (bends)."""
    n = len(pos)
    order = [0]
    left = np.ones(n, dtype=bool)
    left[0] = False
    cur = 0
    for _ in range(n - 1):
        d = np.sum((pos - pos[cur]) ** 2, axis=1)
        d[~left] = np.inf
        cur = int(np.argmin(d))
        left[cur] = False
        order.append(cur)
    order = np.array(order)
    bonds = np.stack([order[:-1], order[1:]])
    bends = np.stack([order[:-2], order[2:]])
    return bonds, bends


def _lj_tables(sig, eps):
    """Packed ecc12/ecc6 with Lorentz-Berthelot mixing over per-type sigma/eps."""
    nt = len(sig)
    i, j = np.triu_indices(nt)          # i <= j (0-based)
    ij = (j + 1) * j // 2 + i           # 0-based packed index of (i+1, j+1)
    s = 0.5 * (sig[i] + sig[j])
    e = np.sqrt(eps[i] * eps[j])
    e12, e6 = lj_internal(e, s)
    ecc12 = np.zeros(nt * (nt + 1) // 2)
    ecc6 = np.zeros_like(ecc12)
    ecc12[ij] = e12
    ecc6[ij] = e6
    tt = np.zeros((nt, nt), dtype=np.int32)
    tt[i, j] = ij + 1
    tt[j, i] = ij + 1
    return np.asfortranarray(tt), ecc12, ecc6


def _assemble(L, solute_pos, solute_q, solute_sig, solute_eps, water_sites, *, alpha, nfft, order, rcut,
              solute_bonds=None, solute_bends=None, solute_mass=12.011, name="", solute_types=None):
    """Build a System.  Solute LJ: one type per solute atom (sigma/eps per atom),
    or, with solute_types (0-based type of each solute atom), sigma/eps per type."""
    ns = len(solute_pos)
    nw = len(water_sites) // 3
    n = ns + 3 * nw
    cart = np.concatenate([solute_pos.reshape(-1, 3), water_sites.reshape(-1, 3)])
    half = L / 2.0
    co = np.asfortranarray(np.diag([half, half, half]))
    oc = np.asfortranarray(np.diag([1.0 / half] * 3))
    xp = cart / half
    q = np.concatenate([np.asarray(solute_q, float), np.tile(WATER_Q, nw)])
    mass = np.concatenate([np.full(ns, solute_mass), np.tile(WATER_MASS, nw)])
    # types: the solute's (one per atom unless solute_types is given), then O, H
    nst = ns if solute_types is None else len(solute_sig)
    sig = np.concatenate([np.asarray(solute_sig, float), [O_SIGMA, 1.0]])
    eps = np.concatenate([np.asarray(solute_eps, float), [O_EPS, 0.0]])
    tt, ecc12, ecc6 = _lj_tables(sig, eps)
    st = np.arange(1, ns + 1) if solute_types is None else np.asarray(solute_types) + 1
    nbtype = np.concatenate([st, np.tile([nst + 1, nst + 2, nst + 2], nw)]).astype(np.int32)
    # groups: each solute atom its own group, each water one group
    ng = ns + nw
    first = np.concatenate([np.arange(ns), ns + 3 * np.arange(nw)])
    last = np.concatenate([np.arange(ns), ns + 3 * np.arange(nw) + 2])
    grppt = np.asfortranarray(np.stack([first + 1, last + 1]).astype(np.int32))
    atomg = np.concatenate([np.arange(1, ns + 1), ns + 1 + np.repeat(np.arange(nw), 3)]).astype(np.int32)
    nsolmol = 1 if ns else 0
    atomp = np.concatenate([np.ones(ns, dtype=np.int32),
                            nsolmol + 1 + np.repeat(np.arange(nw), 3)]).astype(np.int32)
    ss = np.concatenate([np.ones(ns, dtype=np.int32), np.full(3 * nw, 2, dtype=np.int32)])
    gmass_tot = np.zeros(ng)
    np.add.at(gmass_tot, atomg - 1, mass)
    pmass = mass / gmass_tot[atomg - 1]
    w0 = ns + 3 * np.arange(nw)
    cons = np.concatenate([np.stack([w0, w0 + 1]), np.stack([w0, w0 + 2]), np.stack([w0 + 1, w0 + 2])], axis=1) + 1
    sys = System(L=L, xp0=np.ascontiguousarray(xp[:, 0]), yp0=np.ascontiguousarray(xp[:, 1]),
                 zp0=np.ascontiguousarray(xp[:, 2]), charge=charge_internal(q), mass=mass, nbtype=nbtype,
                 ntypes=len(sig), type_table=tt, ecc12=ecc12, ecc6=ecc6, grppt=grppt, atomg=atomg, atomp=atomp,
                 ss_index=ss, pmass=pmass, co=co, oc=oc, volume=8.0 * half ** 3, alpha=alpha,
                 nfft=tuple(nfft), order=order, rcut=rcut, constraints=cons.astype(np.int32),
                 bonds=None if solute_bonds is None else (solute_bonds + 1).astype(np.int32),
                 bends=None if solute_bends is None else (solute_bends + 1).astype(np.int32), name=name)
    return sys


def water_box(seed=1000, L=18.64, nside=6, jitter=0.3, alpha=0.40, nfft=(20, 20, 20), order=5, rcut=8.5):
    """216 rigid waters on a jittered 6^3 lattice in an 18.64 A box (the small CPU-runnable case)."""
    rng = np.random.default_rng(seed)
    a = L / nside
    g = (np.arange(nside) + 0.5) * a - L / 2.0
    o = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    o = o + rng.uniform(-jitter, jitter, size=o.shape)
    sites = _place_waters(rng, o)
    return _assemble(L, np.zeros((0, 3)), [], [], [], sites, alpha=alpha, nfft=nfft, order=order, rcut=rcut,
                     name=f"water{len(o)}")


@functools.lru_cache(maxsize=8)
def _solute(seed, n_solute, radius):
    """Solute atoms of one seed: positions, charges (net 0), LJ sigma/eps, chain exclusions."""
    rng_s = np.random.default_rng(seed)
    pos = _rsa_sphere(rng_s, n_solute, radius, 1.2)
    q = rng_s.normal(0.0, 0.35, n_solute)
    q = q - q.mean()
    sig = rng_s.uniform(2.0, 3.8, n_solute)
    eps = rng_s.uniform(0.02, 0.2, n_solute)
    bonds, bends = _chain_exclusions(pos)
    for a in (pos, q, sig, eps, bonds, bends):
        a.setflags(write=False)
    return pos, q, sig, eps, bonds, bends


def solvated_solute(seed=1001, L=61.0, n_solute=1660, radius=16.0, n_water=7600, alpha=0.35,
                    nfft=(64, 64, 64), order=5, rcut=10.0, charge_scale=1.0, eps_scale=1.0, solute_seed=None):
    """Compact random solute (1-2/1-3 chain exclusions) in a box of rigid waters.

    Default sizes: 1,660 solute atoms + 7,600 waters = 24,460 atoms, L = 61 A,
    grid 64^3, order 5, alpha 0.35, cutoff 10 A.  ``charge_scale`` and
    ``eps_scale`` scale the solute charges and LJ well depths (Hamiltonian
    replica scaling: charges by sqrt(s), epsilon by s).  ``solute_seed``
    fixes the solute independently of the water seed.
    """
    pos, q, sig, eps, bonds, bends = _solute(seed if solute_seed is None else solute_seed, n_solute, radius)
    rng = np.random.default_rng(seed)
    if solute_seed is not None:
        rng = np.random.default_rng(seed + 7)
    # water oxygens on the finest cubic lattice that leaves room for n_water sites
    nside = int(math.ceil(L / 3.1067))
    while True:
        a = L / nside
        g = (np.arange(nside) + 0.5) * a - L / 2.0
        o = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
        far = np.sqrt(np.sum(o * o, axis=1)) > radius + 1.5
        if far.sum() >= n_water:
            break
        nside += 1
    cand = o[far]
    pick = np.sort(rng.choice(len(cand), n_water, replace=False))
    o = cand[pick] + rng.uniform(-0.3, 0.3, size=(n_water, 3))
    sites = _place_waters(rng, o)
    sys = _assemble(L, pos, q * math.sqrt(charge_scale), sig, eps * eps_scale, sites, alpha=alpha, nfft=nfft,
                    order=order, rcut=rcut, solute_bonds=bonds, solute_bends=bends,
                    name=f"solute{n_solute}+water{n_water}")
    return sys


def config(name, seed=None):
    """Named workloads: 'water216' (648 atoms) and 'solvated24k' (24,460 atoms)."""
    if name == "water216":
        return water_box(seed=1000 if seed is None else seed)
    if name == "solvated24k":
        return solvated_solute(seed=1001 if seed is None else seed)
    raise KeyError(name)


def hrem_scales(n=32, s_min=0.3):
    """Hamiltonian scaling factors s_r, geometric from 1 down to s_min."""
    if n == 1:
        return np.ones(1)
    return s_min ** (np.arange(n) / (n - 1))


def hrem_replicas(n=32, s_min=0.3, solute_seed=1001, water_seed=1002, **kw):
    """HREM replicas of one solvated solute: replica r scales the solute charges
    by sqrt(s_r) and the solute LJ epsilon by s_r (s geometric from 1 to s_min)
    and has its own water placement (seed water_seed + 7919 r).  Defaults are
    the 24,460-atom system; keyword arguments go to ``solvated_solute``."""
    return [solvated_solute(seed=water_seed + 7919 * r, solute_seed=solute_seed, charge_scale=float(sr),
                            eps_scale=float(sr), **kw) for r, sr in enumerate(hrem_scales(n, s_min))]


def small_solvated(seed=5, L=24.0, n_solute=40, radius=5.0, n_water=380, nfft=(24, 24, 24), rcut=9.0,
                   alpha=0.35, order=5):
    """A small solute-in-water box for parity tests that exercise solute exclusions."""
    return solvated_solute(seed=seed, L=L, n_solute=n_solute, radius=radius, n_water=n_water, alpha=alpha,
                           nfft=nfft, order=order, rcut=rcut)


# ---- alchemy ---------------------------------------------------------------

@dataclasses.dataclass
class Alchemy:
    """Alchemical description of a System, in force_alchemy.f's terms.

    ss_alchemy: per atom 1 (normal), 2 (added), -2 (removed) (cpropar,
    initialize_lambda.f:35,49).  rmin, epsm: per-atom soft-core LJ parameters,
    sigma_ij = rmin_i + rmin_j (2 sqrt(rmin_i rmin_j) for OPLS) and
    epsmin = 4 sqrt(epsm_i epsm_j) in internal energy units
    (force_alchemy.f:431-437).  qsoft, ljsoft: Beutler soft-core parameters
    (defaults 0.5, blkdta.f90:477-478); intramol, ljintra: the INTRAMOL switches
    of alch_set (defaults false, blkdta.f90:473-474).  shifted selects the
    shifted potential (force_alchemy_shifted.f, SHIFT-POT) with qshift, ljshift
    (defaults 4.0, 0.35, blkdta.f90:479-480) in place of the soft core
    (mts_forces.f:130).  listqq (3, n), when set, is the intra-alchemical
    pair list of ferrf_alchemy_add (alchemy_qq_list).
    """
    ss_alchemy: np.ndarray
    rmin: np.ndarray
    epsm: np.ndarray
    qsoft: float = 0.5
    ljsoft: float = 0.5
    intramol: bool = False
    ljintra: bool = False
    opls: bool = False
    shifted: bool = False
    qshift: float = 4.0
    ljshift: float = 0.35
    listqq: np.ndarray = None

    def soft_params(self):
        """(q, lj) parameters of the selected form."""
        return (self.qshift, self.ljshift) if self.shifted else (self.qsoft, self.ljsoft)


def lambda_state(lrm, lrm0=None, lrmq=None, lrmq0=None, ladd=1.0, ladd0=None, laddq=None, laddq0=None):
    """The eight lambdas force_alchemy.f reads, in the order the library and
    the oracle take them: ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0.
    Unset lambda0 values equal lambda; unset charge lambdas (and their
    lambda0) equal the LJ ones."""
    lrm0 = lrm if lrm0 is None else lrm0
    ladd0 = ladd if ladd0 is None else ladd0
    if lrmq0 is None:
        lrmq0 = lrm0 if lrmq is None else lrmq
    if laddq0 is None:
        laddq0 = ladd0 if laddq is None else laddq
    lrmq = lrm if lrmq is None else lrmq
    laddq = ladd if laddq is None else laddq
    return np.array([ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0], dtype=np.float64)


def alchemy_charges(charge, ss_alchemy, laddq, lrmq):
    """Charges the reciprocal sum and the corrections see at (laddq, lrmq):
    per-atom lambdaq = laddq (added), -lrmq (removed), 0 (normal)
    (mtsmd.f:748-757), charge (1 - |lambdaq|) q with the end points exact
    (mts_furier_alchemy.f:98-111)."""
    lq = np.where(ss_alchemy == 2, laddq, np.where(ss_alchemy == -2, -lrmq, 0.0)).astype(np.float64)
    a = np.abs(lq)
    return np.where(a == 1.0, 0.0, np.where(a == 0.0, charge, (1.0 - a) * charge))


def alchemy_lambdas(ss_alchemy, lam):
    """Per-atom (lambda, lambdaq, lambda0, lambdaq0) of the eight-lambda state
    lam (ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0): added atoms carry
    +ladd, removed atoms -lrm, the others 0 (mtsmd.f:746-757, 2976-2991)."""
    ss = np.asarray(ss_alchemy)
    lam = np.asarray(lam, dtype=np.float64)

    def per(a, r):
        return np.where(ss == 2, a, np.where(ss == -2, -r, 0.0)).astype(np.float64)
    return per(lam[0], lam[1]), per(lam[4], lam[5]), per(lam[2], lam[3]), per(lam[6], lam[7])


def alchemy_qq_list(added, removed, excluded_pairs=()):
    """The intra-alchemical pair list of initialize_lambda.f:66-140: every pair
    of the added atoms followed by the removed atoms (1-based), in that order,
    minus the pairs listed as bonds, constraints, 1-3 or 1-4 (excluded_pairs:
    arrays (2, n), 1-based); facQ is 1 when both atoms are added or both
    removed, else 0.  initialize_lambda.f:135 looks the kinds up by atom
    number in an array filled by position; this builder takes each atom's
    kind by position, which agrees with the reference whenever the
    alchemical atoms are atoms 1..n of the topology (the synthetic ligands)."""
    atoms = [int(a) for a in added] + [int(a) for a in removed]
    kind = [1] * len(added) + [0] * len(removed)
    ex = set()
    for arr in excluded_pairs:
        arr = np.asarray(arr)
        for a, b in zip(arr[0], arr[1]):
            ex.add((int(a), int(b)))
            ex.add((int(b), int(a)))
    out = []
    for i in range(len(atoms) - 1):
        for j in range(i + 1, len(atoms)):
            if (atoms[i], atoms[j]) in ex:
                continue
            out.append((atoms[i], atoms[j], 1 if kind[i] == kind[j] else 0))
    return np.array(out, dtype=np.int32).T.reshape(3, -1)


def alchemy_for(s, alchemical_atoms, removed=True, **kw):
    """An Alchemy for System s whose atoms alchemical_atoms (0-based) are
    removed (or added): rmin = sigma / 2 and epsm = eps (internal units) of
    each atom's LJ type, so that at full coupling the soft-core pair is the
    system's own Lorentz-Berthelot LJ pair; atoms without LJ get rmin 0.1
    (mtsmd.f:1300-1303)."""
    n = s.natoms
    ss = np.ones(n, dtype=np.int32)
    ss[np.asarray(alchemical_atoms, dtype=np.int64)] = -2 if removed else 2
    t = s.nbtype - 1
    lij = s.type_table[t, t] - 1                      # diagonal (i, i) entries
    c12, c6 = s.ecc12[lij], s.ecc6[lij]
    with np.errstate(divide="ignore", invalid="ignore"):
        sig = np.where(c6 > 0, (c12 / np.where(c6 > 0, c6, 1.0)) ** (1.0 / 6.0), 0.0)
        eps4 = np.where(c12 > 0, c6 * c6 / np.where(c12 > 0, c12, 1.0), 0.0)   # 4 eps
    epsm = eps4 / 4.0
    rmin = np.where(epsm == 0.0, 0.1, sig / 2.0)
    return Alchemy(ss_alchemy=ss, rmin=rmin, epsm=epsm, **kw)


def alchemy_windows(n_windows=16, seed=3001, L=40.0, n_ligand=48, radius=5.5, n_water=2117, alpha=0.35,
                    nfft=(48, 48, 48), order=5, rcut=9.0):
    """lambda-windows of one ligand decoupled in water: a 48-atom ligand
    (neutral charges ~N(0, 0.3) e, random spanning-chain exclusions) in a
    40 A box of rigid waters (6,399 atoms), grid 48^3, order 5, cutoff 9 A.
    Window w removes the ligand at lambda_w = w / (n - 1) for both LJ and
    charge; its lambda0 is the next window's (the last window's, the one
    before), so each window's dwrk is the energy step to its neighbour.
    Returns (system, alchemy, [lambda_state per window])."""
    rng = np.random.default_rng(seed)
    pos = _rsa_sphere(rng, n_ligand, radius, 1.2)
    q = rng.normal(0.0, 0.3, n_ligand)
    q = q - q.mean()
    sig = rng.uniform(2.0, 3.8, n_ligand)
    eps = rng.uniform(0.02, 0.2, n_ligand)
    bonds, bends = _chain_exclusions(pos)
    nside = int(math.ceil(L / 3.1067))
    while True:
        a = L / nside
        g = (np.arange(nside) + 0.5) * a - L / 2.0
        o = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
        far = np.sqrt(np.sum(o * o, axis=1)) > radius + 1.5
        if far.sum() >= n_water:
            break
        nside += 1
    cand = o[far]
    pick = np.sort(rng.choice(len(cand), n_water, replace=False))
    o = cand[pick] + rng.uniform(-0.3, 0.3, size=(n_water, 3))
    sites = _place_waters(rng, o)
    s = _assemble(L, pos, q, sig, eps, sites, alpha=alpha, nfft=nfft, order=order, rcut=rcut, solute_bonds=bonds,
                  solute_bends=bends, name=f"ligand{n_ligand}+water{n_water}")
    al = alchemy_for(s, np.arange(n_ligand))
    al.listqq = alchemy_qq_list([], np.arange(1, n_ligand + 1), (s.bonds, s.bends))
    lam = np.linspace(0.0, 1.0, n_windows) if n_windows > 1 else np.zeros(1)
    states = []
    for w in range(n_windows):
        l0 = lam[w + 1] if w + 1 < n_windows else (lam[w - 1] if w > 0 else lam[w])
        states.append(lambda_state(float(lam[w]), lrm0=float(l0)))
    return s, al, states
