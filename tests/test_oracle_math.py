"""Pins of the oracle by mathematics, independent of the reference build.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a transposed operand, an off-by-one grid offset) fails.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import rel, relrms


def cardinal_bspline(n, x):
    """M_n(x) = 1/(n-1)! sum_k (-1)^k C(n,k) (x-k)_+^(n-1)  (closed form)."""
    s = 0.0
    for k in range(n + 1):
        if x - k > 0:
            s += (-1) ** k * math.comb(n, k) * (x - k) ** (n - 1)
    return s / math.factorial(n - 1)


@pytest.mark.parametrize("order", [3, 4, 5, 6, 8])
def test_bspline_closed_form(order):
    for w in (0.0, 0.137, 0.5, 0.91, 0.999):
        th, dth = O.fill_bspline(w, order)
        # theta(i) = M_n(w + n - i), dtheta(i) = M_n'(w + n - i) = M_{n-1}(x) - M_{n-1}(x-1)
        for i in range(1, order + 1):
            x = w + order - i
            # the alternating closed-form sum itself cancels to ~1e-14 near x = n
            assert abs(th[i - 1] - cardinal_bspline(order, x)) < 1e-12
            d = cardinal_bspline(order - 1, x) - cardinal_bspline(order - 1, x - 1)
            assert abs(dth[i - 1] - d) < 1e-12
        assert abs(th.sum() - 1.0) < 1e-15 and abs(dth.sum()) < 1e-15


@pytest.mark.parametrize("order,K", [(4, 20), (5, 20), (5, 64), (6, 48), (7, 30)])
def test_bsp_moduli_closed_form(order, K):
    """|sum_j M_n(n-j) e^{2 pi i k j / K}|^2 with 2*pi from the REAL(4) pi
    literal (pme_fftw3.f90:214), values below 1e-7 replaced by the mean of
    their neighbours (pme_fftw3.f90:226-228)."""
    twopi32 = 2.0 * float(np.float32(np.pi))
    arr = np.array([cardinal_bspline(order, order - j) for j in range(0, order)])  # M_n(n), ..., M_n(1)
    b = np.zeros(K)
    b[1:order + 1] = arr[::-1][::-1] if False else [cardinal_bspline(order, order - (j - 1)) for j in range(1, order + 1)]
    k = np.arange(K)
    j = np.arange(K)
    arg = twopi32 * np.outer(k, j) / K
    mod = (b * np.cos(arg)).sum(1) ** 2 + (b * np.sin(arg)).sum(1) ** 2
    for i in range(K):
        if mod[i] < 1e-7:
            mod[i] = 0.5 * (mod[i - 1] + mod[i + 1])
    got = O.bsp_moduli(order, K)
    assert np.max(np.abs(got - mod) / mod) < 1e-13
    if order % 2 == 1 and K % 2 == 0:
        assert abs(got[K // 2] - 0.5 * (got[K // 2 - 1] + got[K // 2 + 1])) < 1e-15  # the Nyquist fix-up fired
    # the REAL(4) pi breaks the mirror symmetry |b(m)| = |b(-m)| at ~1e-7
    asym = np.max(np.abs(got[1:] - got[1:][::-1]) / got[1:])
    assert 1e-10 < asym < 1e-5


@pytest.mark.parametrize("shape", [(20, 20, 20), (24, 20, 30), (16, 12, 10), (48, 48, 48)])
def test_fft_matches_pocketfft(shape):
    rng = np.random.default_rng(0)
    Q = rng.normal(size=shape[::-1])
    F = O.fft_r2c(Q)
    assert np.abs(F - np.fft.rfftn(Q)).max() < 1e-13 * np.abs(F).max()
    back = O.fft_c2r(F, shape[0])
    assert np.abs(back - Q * np.prod(shape)).max() < 1e-12 * np.prod(shape)


def _two_ions(L=20.0, K=48, order=8, alpha=0.3):
    import dataclasses
    from orac_hip import synth
    s = synth.water_box()
    half = L / 2
    co = np.asfortranarray(np.diag([half] * 3))
    oc = np.asfortranarray(np.diag([1 / half] * 3))
    xp = np.array([0.05, 0.31, -0.42]) / 1.0
    yp = np.array([-0.2, 0.44, 0.13])
    zp = np.array([0.7, -0.33, 0.02])
    q = np.array([0.8, -0.5, -0.3])
    return dataclasses.replace(s, L=L, xp0=xp, yp0=yp, zp0=zp, charge=q, mass=np.ones(3), nbtype=np.ones(3, np.int32),
                               atomp=np.array([1, 2, 3], np.int32), atomg=np.array([1, 2, 3], np.int32),
                               grppt=np.asfortranarray(np.array([[1, 2, 3], [1, 2, 3]], np.int32)),
                               ss_index=np.full(3, 2, np.int32), pmass=np.ones(3), co=co, oc=oc,
                               volume=L ** 3, alpha=alpha, nfft=(K, K, K), order=order, constraints=None)


def ewald_kspace(s, mmax=12):
    """Explicit reciprocal Ewald sum E = 1/(2 pi V) sum_{m != 0} exp(-pi^2 m^2/a^2)/m^2 |S(m)|^2,
    with the REAL(4) pi the reference uses in scalar_sum.f:18."""
    pi = float(np.float32(np.pi))
    x, y, z = s.cartesian()
    r = np.stack([x, y, z], 1)
    L = s.L
    e = 0.0
    rng = np.arange(-mmax, mmax + 1)
    M = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    M = M[np.any(M != 0, axis=1)] / L
    m2 = np.sum(M * M, 1)
    S = np.exp(2j * np.pi * (M @ r.T)) @ s.charge
    e = np.sum(np.exp(-pi * pi * m2 / s.alpha ** 2) / m2 * np.abs(S) ** 2) / (2 * pi * s.volume)
    return e


def test_reciprocal_energy_converges_to_ewald_sum():
    s = _two_ions()
    e_pme = O.fft_pme(s, pressure=False)["eer"]
    e_ew = ewald_kspace(s)
    assert rel(e_pme, e_ew) < 2e-6


def test_reciprocal_force_is_minus_gradient():
    s = _two_ions(K=32, order=6)
    base = O.fft_pme(s, pressure=False)
    h = 1e-6
    for d, name in enumerate(("xp0", "yp0", "zp0")):
        a = 1
        c = getattr(s, name).copy()
        c[a] += h / (s.L / 2)
        sp = s.with_coords(*(c if k == d else getattr(s, n) for k, n in enumerate(("xp0", "yp0", "zp0"))))
        c2 = getattr(s, name).copy()
        c2[a] -= h / (s.L / 2)
        sm = s.with_coords(*(c2 if k == d else getattr(s, n) for k, n in enumerate(("xp0", "yp0", "zp0"))))
        fd = -(O.fft_pme(sp, pressure=False)["eer"] - O.fft_pme(sm, pressure=False)["eer"]) / (2 * h)
        assert abs(fd - base["f"][d, a]) < 1e-6 * np.abs(base["f"]).max()


@pytest.mark.parametrize("variant", ["fftw", "fftpack"])
def test_reciprocal_virial_is_strain_derivative(variant):
    """vir(a,b) = dE/d(eps_ab) at fixed fractional coordinates.  The FFTW
    branch's fac12/fac13/fac23 masks (scalar_sum.f:63-115) act only where
    k = K/2 aliases +-K/2, whose terms are negligible here, so both branches
    match the strain derivative (the masks themselves are pinned against the
    compiled reference's scalar_sum in test_oracle_vs_reference)."""
    import dataclasses
    from orac_hip import synth
    s = synth.water_box(nfft=(24, 24, 24))
    h = 1e-6

    def energy(eps):
        co = np.asfortranarray((np.eye(3) + eps) @ s.co)
        oc = np.asfortranarray(np.linalg.inv(co))
        s2 = dataclasses.replace(s, co=co, oc=oc, volume=8 * np.linalg.det(co))
        return O.fft_pme(s2, pressure=False, variant=variant)["eer"]

    # reciprocal virial alone: stressc = -vir + molecular term -> isolate vir via pressure=True minus correction
    x, y, z = s.cartesian()
    recip = O.recip_of(s.oc)
    g = O.charge_grid(x, y, z, s.charge, recip, s.nfft, s.order)
    bsp = [O.bsp_moduli(s.order, k) for k in s.nfft]
    if variant == "fftw":
        _, vir, _ = O.scalar_sum(np.fft.rfftn(g["Q"]), s.alpha, s.volume, recip, bsp, s.nfft)
    else:
        import ctypes as C
        Qc = np.ascontiguousarray(np.fft.fftn(g["Q"]))
        v = np.zeros(9)
        e = C.c_double()
        O.lib().orc_scalar_sum_full(O._p(Qc), C.c_double(s.alpha), C.c_double(s.volume), O._p(O._f(recip)),
                                    O._p(bsp[0]), O._p(bsp[1]), O._p(bsp[2]), C.c_int(24), C.c_int(24), C.c_int(24),
                                    C.c_double(1e9), C.byref(e), O._p(v))
        vir = v.reshape(3, 3, order="F")
    for a in range(3):
        eps = np.zeros((3, 3))
        eps[a, a] = h
        fd = (energy(eps) - energy(-eps)) / (2 * h)
        assert abs(fd - vir[a, a]) < 1e-6 * abs(vir[a, a]) + 1e-12
    eps = np.zeros((3, 3))
    eps[0, 1] = eps[1, 0] = h / 2
    fd = (energy(eps) - energy(-eps)) / (2 * h)
    assert abs(fd - vir[0, 1]) < 1e-5 * np.abs(vir).max()


def brute_pairs(s, rcut):
    """O(Ng^2) minimum-image group-pair set with the reference's full-co distance."""
    xpg = s.group_centres()
    ng = s.ngrp
    out = set()
    for n in range(ng):
        d = xpg[:, n:n + 1] - xpg[:, n + 1:]
        d = d - 2.0 * np.round(0.5 * d)
        c = s.co @ d
        r2 = np.sum(c * c, 0)
        for l in np.nonzero(r2 < rcut * rcut)[0]:
            out.add((n + 1, n + 2 + int(l)))
    return out


def list_pairs(nnlpp, ng):
    out, pos = set(), 0
    for n in range(1, ng + 1):
        m = nnlpp[pos]
        for l in nnlpp[pos + 1: pos + 1 + m]:
            out.add((n, int(l)))
        pos += m + 1
    return out


def test_linked_cell_list_equals_brute_force(small_solute):
    s = small_solute
    rc = s.rcut + s.rtol + s.rnei
    nc = int(s.L // 3.0)
    fa = O.ForceArgs(s)
    st = O.lc_index(nc, nc, nc, rc, s.co)
    lst = O.lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
    assert list_pairs(lst, s.ngrp) == brute_pairs(s, rc)


def test_lj_force_is_minus_gradient_with_switching(small_solute):
    """Zero charges: F = -grad(E) holds exactly for LJ including the group
    switch S(R) and its S'V term spread by mass fraction (force.f:756-817)."""
    import dataclasses
    s = dataclasses.replace(small_solute, charge=np.zeros(small_solute.natoms))
    rc = s.rcut + s.rtol + s.rnei
    nnlpp = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0)[
        "nnlpp1"]
    kw = dict(rinn=3.0, rout=6.0, rtolinn=0.6, rtolout=0.8, rneigh=0.0)

    def energy(s2):
        r = O.force(s2, nnlpp, **kw)
        return r["uconf_slt"] + r["uconf_slv"] + r["uconf_ss"], r["f"]

    e0, f0 = energy(s)
    rng = np.random.default_rng(4)
    h = 1e-6
    for a in rng.choice(s.natoms, 6, replace=False):
        for d, name in enumerate(("xp0", "yp0", "zp0")):
            cs = [s.xp0.copy(), s.yp0.copy(), s.zp0.copy()]
            cs[d][a] += h / (s.L / 2)
            ep, _ = energy(s.with_coords(*cs))
            cs[d][a] -= 2 * h / (s.L / 2)
            em, _ = energy(s.with_coords(*cs))
            fd = -(ep - em) / (2 * h)
            assert abs(fd - f0[d, a]) < 1e-6 * np.abs(f0).max() + 1e-9, (a, d, fd, f0[d, a])


def test_direct_newton_third_law(small_solute):
    s = small_solute
    rc = s.rcut + s.rtol + s.rnei
    nnlpp = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0)[
        "nnlpp1"]
    r = O.force(s, nnlpp, rinn=2.5, rout=s.rcut, rtolinn=0.5, rtolout=0.3, rneigh=0.0)
    assert np.abs(r["f"].sum(axis=1)).max() < 1e-12 * np.abs(r["f"]).max()


def test_direct_energy_matches_brute_force(water216):
    """Every O-O and O-H/H-H pair of distinct waters whose centres are within
    rcut: A&S erfc Coulomb + LJ, summed without lists or cells."""
    s = water216
    rc = s.rcut + s.rtol + s.rnei
    nnlpp = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0)[
        "nnlpp1"]
    r = O.force(s, nnlpp, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=0.0, rneigh=0.0)
    x, y, z = s.cartesian()
    pos = np.stack([x, y, z], 1)
    xpg = (s.co @ s.group_centres()).T
    L = s.L
    a1, a2, a3, a4, a5, p = 0.2548296, -0.28449674, 1.4214137, -1.453152, 1.0614054, 0.3275911
    ecoul = elj = 0.0
    c12, c6 = s.ecc12[0], s.ecc6[0]
    for gi in range(s.ngrp):
        for gj in range(gi + 1, s.ngrp):
            dg = xpg[gi] - xpg[gj]
            shift = L * np.round(dg / L)
            if np.sum((dg - shift) ** 2) >= s.rcut ** 2:
                continue
            for ai in range(3 * gi, 3 * gi + 3):
                for aj in range(3 * gj, 3 * gj + 3):
                    d = pos[ai] - pos[aj] - shift
                    rr = math.sqrt(d @ d)
                    t = 1 / (1 + p * s.alpha * rr)
                    erfc = ((((a5 * t + a4) * t + a3) * t + a2) * t + a1) * t * math.exp(-(s.alpha * rr) ** 2)
                    ecoul += s.charge[ai] * s.charge[aj] * erfc / rr
                    if ai % 3 == 0 and aj % 3 == 0:
                        elj += c12 / rr ** 12 - c6 / rr ** 6
    assert rel(r["ucoul_slv"], ecoul) < 1e-11
    assert rel(r["uconf_slv"], elj) < 1e-11


def test_ferrf_and_cself_closed_forms(small_solute):
    s = small_solute
    pair = s.bonds[:, :1]
    ia, ib = pair[0, 0] - 1, pair[1, 0] - 1
    x, y, z = s.cartesian()
    d = np.array([x[ia] - x[ib], y[ia] - y[ib], z[ia] - z[ib]])
    r = math.sqrt(d @ d)
    ar = s.alpha * r
    t = 1 / (1 + 0.3275911 * ar)
    erfc = ((((1.0614054 * t - 1.453152) * t + 1.4214137) * t - 0.28449674) * t + 0.2548296) * t * math.exp(-ar * ar)
    e = -s.charge[ia] * s.charge[ib] * (1 - erfc) / r
    res = O.ferrf(s, pair, 1, 1.0)
    assert rel(res["slt"], e) < 1e-14
    # force = +q q (erf - 2/sqrt(pi) a r e^{-a^2 r^2}) / r^3 * d on ia, opposite on ib
    fmag = -s.charge[ia] * s.charge[ib] * ((1 - erfc) - 2 / math.sqrt(math.pi) * ar * math.exp(-ar * ar)) / r ** 3
    assert np.allclose(res["f"][:, ia], fmag * d, rtol=1e-13, atol=0)
    assert np.allclose(res["f"][:, ib], -fmag * d, rtol=1e-13, atol=0)
    slt, slv = O.cself(s)
    q = s.charge
    assert rel(slv, -np.sum(q[s.ss_index == 2] ** 2) * s.alpha / math.sqrt(math.pi)) < 1e-13
    assert rel(slt, -np.sum(q[s.ss_index == 1] ** 2) * s.alpha / math.sqrt(math.pi)) < 1e-13
