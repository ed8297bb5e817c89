"""The oracle's bonded terms (fptors, fnb14) against what the mathematics
fixes, on small synthetic geometries: forces are minus the energy gradient
(central differences) for every multiplicity fptors knows, the energy is
invariant under rigid rotation and the forces of each term sum to zero; the
1-4 pair is the closed form lj_fudge (c12/r^12 - c6/r^6) + fudge q q
erfc(alpha r)/r (A&S 7.1.26) or plain Coulomb.  Bit-exact agreement with the
compiled reference is in tests/test_oracle_fkbp12.py."""
import dataclasses
import math

import numpy as np
import pytest

import oracle as O
from orac_hip import synth

MULTS = [1, 2, 3, 4, 5, 6, 7, 8, 12, 14]


def _system(n, rng, scale=1.5):
    s = synth.water_box()
    pos = rng.normal(0.0, scale, (n, 3))
    half = s.co[0, 0]
    return dataclasses.replace(s, xp0=pos[:, 0] / half, yp0=pos[:, 1] / half, zp0=pos[:, 2] / half,
                               ss_index=np.ones(s.natoms, np.int32)), pos


def _energy(s, pos, ltor, pota, potb):
    n = s.natoms
    x = np.zeros(n); y = np.zeros(n); z = np.zeros(n)
    x[:len(pos)], y[:len(pos)], z[:len(pos)] = pos[:, 0], pos[:, 1], pos[:, 2]
    r = O.fptors(s, ltor, pota, potb, coords=(x, y, z))
    assert r["status"] == 0
    return r["utors_slt"], r["f"][:, :len(pos)]


@pytest.mark.parametrize("mult", MULTS)
@pytest.mark.parametrize("sign", [1, -1])
def test_torsion_force_is_minus_gradient(mult, sign):
    rng = np.random.default_rng(100 + mult)
    s, pos = _system(4, rng)
    ltor = np.array([[1], [2], [3], [4]], np.int32)
    pota, potb = np.array([float(sign * mult)]), np.array([1.7])
    e0, f = _energy(s, pos, ltor, pota, potb)
    h = 1e-6
    for a in range(4):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[a, c] += h
            p2[a, c] -= h
            fd = -(_energy(s, p1, ltor, pota, potb)[0] - _energy(s, p2, ltor, pota, potb)[0]) / (2 * h)
            assert abs(fd - f[c, a]) <= 1e-6 * max(1.0, np.abs(f).max()), (mult, sign, a, c)
    assert np.abs(f.sum(axis=1)).max() <= 1e-12 * np.abs(f).max()


def test_torsion_energy_is_rotation_invariant():
    rng = np.random.default_rng(7)
    s, pos = _system(12, rng)
    ltor = np.array([[1, 5, 9], [2, 6, 10], [3, 7, 11], [4, 8, 12]], np.int32)
    pota, potb = np.array([3.0, -2.0, 1.0]), np.array([1.1, 0.4, 2.0])
    e0, _ = _energy(s, pos, ltor, pota, potb)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    e1, _ = _energy(s, pos @ q.T + 0.3, ltor, pota, potb)
    assert abs(e1 - e0) <= 1e-12 * abs(e0)


def test_one_four_pair_closed_form():
    s = synth.small_solvated()
    i14 = np.array([[1, 3], [2, 7]], np.int32)  # pairs (1,2) and (3,7)
    t14 = np.array([1, 4], np.int32)
    x, y, z = s.cartesian()
    for ewald in (True, False):
        r = O.fnb14(s, i14, t14, 0.8, 0.5, ewald=ewald, cut=True)
        e_lj = e_q = 0.0
        for k in range(2):
            a, b = i14[0, k] - 1, i14[1, k] - 1
            d = math.sqrt((x[a] - x[b]) ** 2 + (y[a] - y[b]) ** 2 + (z[a] - z[b]) ** 2)
            c12, c6 = s.ecc12[t14[k] - 1], s.ecc6[t14[k] - 1]
            e_lj += 0.5 * (c12 / d ** 12 - c6 / d ** 6)
            ar = s.alpha * d
            t = 1.0 / (1.0 + 0.3275911 * ar)
            erfc = ((((1.0614054 * t - 1.453152) * t + 1.4214137) * t - 0.28449674) * t + 0.2548296) * t * \
                math.exp(-ar * ar)
            e_q += 0.8 * s.charge[a] * s.charge[b] * (erfc if ewald else 1.0) / d
        assert abs(r["uconf_slt"] - e_lj) <= 1e-12 * abs(e_lj)
        assert abs(r["ucoul_slt"] - e_q) <= 1e-12 * abs(e_q)
        assert np.abs(r["f"].sum(axis=1)).max() <= 1e-12 * np.abs(r["f"]).max()
