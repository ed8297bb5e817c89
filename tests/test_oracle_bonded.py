"""The oracle's bonded terms (fptors, fitors, fnb14, fpbond, fpbend) against what the mathematics
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


def _bb_system(rng, n=3):
    s, pos = _system(n, rng, scale=1.0)
    return s, pos


def _coords(s, pos):
    n = s.natoms
    x = np.zeros(n); y = np.zeros(n); z = np.zeros(n)
    x[:len(pos)], y[:len(pos)], z[:len(pos)] = pos[:, 0], pos[:, 1], pos[:, 2]
    return x, y, z


def _grad_check(efun, pos, f, h=1e-6, tol=1e-6):
    num = np.zeros_like(pos)
    for a in range(pos.shape[0]):
        for c in range(3):
            p, m = pos.copy(), pos.copy()
            p[a, c] += h
            m[a, c] -= h
            num[a, c] = -(efun(p) - efun(m)) / (2 * h)
    assert np.abs(f.T - num).max() <= tol * max(1.0, np.abs(num).max())
    assert np.abs(f.sum(axis=1)).max() <= 1e-9 * max(1.0, np.abs(f).max())


def test_bond_force_is_minus_gradient_and_closed_form():
    rng = np.random.default_rng(21)
    s, pos = _bb_system(rng, 2)
    lb = np.array([[1], [2]], np.int32)
    r0, k = 1.3, 75.0

    def run(p):
        return O.fpbond(s, lb, [r0], [k], coords=_coords(s, p))

    r = run(pos)
    d = np.linalg.norm(pos[1] - pos[0])
    assert math.isclose(r["ubond_slt"], k * (d - r0) ** 2, rel_tol=1e-14)
    _grad_check(lambda p: run(p)["ubond_slt"], pos, r["f"][:, :2])


@pytest.mark.parametrize("branch", ["plain", "urey_bradley", "coupling_d", "coupling_efg", "linear"])
def test_bend_force_is_minus_gradient_and_closed_form(branch):
    rng = np.random.default_rng({"plain": 1, "urey_bradley": 2, "coupling_d": 3, "coupling_efg": 4,
                                 "linear": 5}[branch])
    s, pos = _bb_system(rng, 3)
    lb = np.array([[1], [2], [3]], np.int32)
    th0, k = 1.9, 60.0
    pots = {"plain": [th0, k, 0, 0, 0, 0, 0], "urey_bradley": [th0, k, 2.4, 30.0, 0, 0, 0],
            "coupling_d": [th0, k, 1.2, 15.0, 1.4, 0, 0], "coupling_efg": [th0, k, 0, 0, 12.0, 1.1, 1.3],
            "linear": [math.pi, k, 0, 0, 0, 0, 0]}[branch]

    def run(p):
        return O.fpbend(s, lb, [[v] for v in pots], coords=_coords(s, p))

    r = run(pos)
    a, b, c = pos
    r12, r32 = np.linalg.norm(a - b), np.linalg.norm(c - b)
    cos = np.dot(a - b, c - b) / (r12 * r32)
    th = math.acos(cos)
    e = 2 * k * (cos + 1) if branch == "linear" else k * (th - th0) ** 2
    if branch == "urey_bradley":
        e += 30.0 * (np.linalg.norm(c - a) - 2.4) ** 2
    if branch == "coupling_d":
        e += 15.0 * (r12 - 1.2) * (r32 - 1.4)
    if branch == "coupling_efg":
        e += 12.0 * (r12 - 1.1) * (r32 - 1.3)
    assert math.isclose(r["ubend_slt"], e, rel_tol=1e-12)
    _grad_check(lambda p: run(p)["ubend_slt"], pos, r["f"][:, :3])


@pytest.mark.parametrize("form", ["harmonic", 1, 2, 3, 4, 6])
def test_improper_force_is_minus_gradient(form):
    """fitors: the harmonic form is potb (phi - pota)^2 on the signed dihedral
    (checked against an independent atan2 dihedral), the cosine forms agree
    with fptors; forces are minus the gradient."""
    rng = np.random.default_rng(31 if form == "harmonic" else 40 + form)
    s, pos = _bb_system(rng, 4)
    lt = np.array([[1], [2], [3], [4]], np.int32)
    if form == "harmonic":
        pa, pb, pc = 0.4, 3.0, 1.0
    else:
        pa, pb, pc = float(form), 1.7, -1.0

    def run(p):
        return O.fitors(s, lt, [pa], [pb], [pc], coords=_coords(s, p))

    r = run(pos)
    assert r["status"] == 0
    if form == "harmonic":
        b1, b2, b3 = pos[1] - pos[0], pos[2] - pos[1], pos[3] - pos[2]
        n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
        phi = math.atan2(np.dot(np.cross(n1, n2), b2 / np.linalg.norm(b2)), np.dot(n1, n2))
        d = (phi - pa + math.pi) % (2 * math.pi) - math.pi
        # the reference wraps the dihedral with 3.14159, not pi: a wrapped
        # dihedral is off by up to 4 (pi - 3.14159) ~ 1e-5 rad
        assert abs(math.sqrt(r["utors_slt"] / pb) - abs(d)) < 2.2e-5
    else:
        t = O.fptors(s, lt, [pa], [pb], coords=_coords(s, pos))
        assert math.isclose(r["utors_slt"], t["utors_slt"], rel_tol=1e-14)
        assert np.allclose(r["f"], t["f"], rtol=0, atol=1e-12 * np.abs(t["f"]).max())
    # the harmonic force divides by sin of the (3.14159-wrapped) dihedral: ~1e-5 relative
    _grad_check(lambda p: run(p)["utors_slt"], pos, r["f"][:, :4], tol=2e-5 if form == "harmonic" else 1e-6)
