"""On-device r-RESPA dynamics of resident systems (the caller side of the path).

``Dynamics`` prepares one ``ResidentSystem`` for integration: rigid 3-site
waters (SHAKE/RATTLE) and, for the solute, harmonic distance terms on its
chain 1-2 and 1-3 pairs (the exclusions the generator draws) with the
generated geometry as the rest lengths.  ``run_batch`` integrates a batch of
such systems on the device, one kernel launch per stage for all of them.

Units are ORAC's internal ones (setuns.f): amu, A, fs, energy amu A^2/fs^2.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import synth
from .api import _check, _f, _i, _p, lib

KB_INTERNAL = 1.380649e-23 / synth.UNITE          # Boltzmann constant, internal energy per K
KCAL_MOL = 4184.0 / synth.EFACT                    # 1 kcal/mol in internal energy units


def rigid_waters(s):
    """First atoms (1-based) and lengths (O-H1, O-H2, H1-H2) of the rigid waters."""
    ss = np.asarray(s.ss_index)
    gp = np.asarray(s.grppt)[:, : s.ngrp]
    solv = ss[gp[0] - 1] == 2
    first = gp[0, solv].astype(np.int32)
    if np.any(gp[1, solv] - gp[0, solv] != 2):
        raise ValueError("solvent groups must be 3-site waters")
    half = np.radians(synth.ANG_HOH) / 2.0
    lens = np.tile([synth.R_OH, synth.R_OH, 2.0 * synth.R_OH * np.sin(half)], (len(first), 1))
    return first, lens


def solute_terms(s, k_bond=300.0, k_bend=50.0):
    """Harmonic terms on the solute chain: 1-2 pairs (k_bond) and 1-3 pairs
    (k_bend), kcal/mol/A^2, rest lengths from the generated coordinates."""
    if s.bonds is None:
        return np.zeros((2, 0), np.int32), np.zeros(0), np.zeros(0)
    pairs = np.concatenate([s.bonds, s.bends], axis=1).astype(np.int32)
    k = np.concatenate([np.full(s.bonds.shape[1], k_bond), np.full(s.bends.shape[1], k_bend)]) * KCAL_MOL
    x = np.stack(s.cartesian())
    d = x[:, pairs[0] - 1] - x[:, pairs[1] - 1]
    d -= s.L * np.round(d / s.L)
    return pairs, k, np.sqrt((d * d).sum(axis=0))


def degrees_of_freedom(s, nrigid):
    return 3 * s.natoms - 3 * nrigid - 3


class Dynamics:
    """Integration state of one resident system."""

    def __init__(self, rs, s, k_bond=300.0, k_bend=50.0):
        self.rs, self.s = rs, s
        first, lens = rigid_waters(s)
        pairs, k, r0 = solute_terms(s, k_bond, k_bend)
        self.nrigid = len(first)
        self.dof = degrees_of_freedom(s, self.nrigid)
        self._keep = (_i(first), _f(lens.ravel()), _i(pairs.T.ravel() if pairs.size else np.zeros(2, np.int32)),
                      _f(k if k.size else np.zeros(1)), _f(r0 if r0.size else np.zeros(1)))
        a, b, c, d, e = self._keep
        _check(lib().orac_hip_md_setup(rs.ctx._h, rs.id, C.c_int32(len(first)), _p(a), _p(b),
                                       C.c_int32(pairs.shape[1]), _p(c), _p(d), _p(e)))

    def set_velocities(self, v):
        v = np.asarray(v, np.float64)
        _check(lib().orac_hip_md_set_state(self.rs.ctx._h, self.rs.id, None, None, None, _p(_f(v[0])),
                                           _p(_f(v[1])), _p(_f(v[2]))))

    def set_positions(self, x):
        x = np.asarray(x, np.float64)
        _check(lib().orac_hip_md_set_state(self.rs.ctx._h, self.rs.id, _p(_f(x[0])), _p(_f(x[1])), _p(_f(x[2])),
                                           None, None, None))

    def maxwell_boltzmann(self, temperature, seed):
        """Velocities at temperature (K), zero total momentum, projected on the constraints."""
        rng = np.random.default_rng(seed)
        m = self.s.mass
        v = rng.normal(size=(3, self.s.natoms)) * np.sqrt(KB_INTERNAL * temperature / m)
        v -= (v * m).sum(axis=1, keepdims=True) / m.sum()
        self.set_velocities(v)

    def state(self):
        n = self.s.natoms
        arr = [np.zeros(n) for _ in range(6)]
        _check(lib().orac_hip_md_get_state(self.rs.ctx._h, self.rs.id, *[_p(a) for a in arr]))
        return np.stack(arr[:3]), np.stack(arr[3:])

    def energies(self):
        """Kinetic, harmonic-term, nonbonded (eer + direct + corrections + self)
        and total energy, temperature, momentum, constraint failures."""
        out = np.zeros(5)
        fail = C.c_int32()
        _check(lib().orac_hip_system_md_energies(self.rs.ctx._h, self.rs.id, _p(out), C.byref(fail)))
        r = self.rs.results()
        nb = (r["eer"] + r["uconf_slt"] + r["uconf_slv"] + r["uconf_ss"] + r["ucoul_slt"] + r["ucoul_slv"]
              + r["ucoul_ss"] + r["corr_slt"] + r["corr_slv"] + r["self_slt"] + r["self_slv"])
        ke, eb = out[0], out[1]
        return dict(kinetic=ke, bonded=eb, nonbonded=nb, total=ke + eb + nb,
                    temperature=2.0 * ke / (self.dof * KB_INTERNAL), momentum=out[2:5], fail=fail.value)


def run_batch(batch, nsteps, dt=2.0, ninner=4, list_every=5, graph=True, nmid=None):
    """nsteps outer steps of dt; with nmid (systems with two shells) the inner
    shell runs nmid times per dt and the fast terms ninner times per dt/nmid."""
    if nmid:
        _check(lib().orac_hip_batch_md_run_shells(batch.ctx._h, batch.id, C.c_int32(nsteps), C.c_double(dt),
                                                  C.c_int32(nmid), C.c_int32(ninner), C.c_int32(list_every),
                                                  C.c_int32(1 if graph else 0)))
        return
    _check(lib().orac_hip_batch_md_run(batch.ctx._h, batch.id, C.c_int32(nsteps), C.c_double(dt),
                                       C.c_int32(ninner), C.c_int32(list_every), C.c_int32(1 if graph else 0)))


def relax_batch(batch, niter, hstep=0.5, dmax=0.02, list_every=5):
    _check(lib().orac_hip_batch_md_relax(batch.ctx._h, batch.id, C.c_int32(niter), C.c_double(hstep),
                                         C.c_double(dmax), C.c_int32(list_every)))


def scale_velocities(batch, scale):
    sc = _f(np.asarray(scale, np.float64))
    _check(lib().orac_hip_batch_md_scale_velocities(batch.ctx._h, batch.id, _p(sc)))


def equilibrate(batch, dyns, temperature, nsteps, every=10, dt=2.0, ninner=4, list_every=5, graph=True):
    """Velocity rescaling to temperature every `every` steps (preparation only)."""
    for _ in range(max(1, nsteps // every)):
        run_batch(batch, every, dt=dt, ninner=ninner, list_every=list_every, graph=graph)
        batch.ctx.sync()
        sc = [np.sqrt(temperature / max(d.energies()["temperature"], 1e-6)) for d in dyns]
        scale_velocities(batch, sc)
