"""The oracle's alchemical routines (force_alchemy, ferrf_alchemy,
cself_alchemy, fft_pme_alchemy) against the compiled reference's results
(tests/golden/alchemy.npz, written by tests/golden/make_alchemy.py) and
against what the mathematics fixes: at full coupling the soft-core pair is
the plain pair (force() of force.f), forces are minus the gradient of the
energies, and dwrk is U(lambda) - U(lambda0) from two single-lambda calls.
"""
import dataclasses
import hashlib
import os
import sys

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_alchemy as G  # noqa: E402
from orac_hip import synth  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden", "alchemy.npz")


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


@pytest.fixture(scope="module")
def ctx():
    s, atoms = G.system()
    fa = O.ForceArgs(s)
    nn = G.outer_list(s, fa)
    g = np.load(GOLD)
    assert np.array_equal(digest(np.concatenate([s.xp0, s.yp0, s.zp0, s.charge])), g["inputs_sha"])
    assert np.array_equal(digest(nn), g["nnlpp_sha"])
    return s, atoms, fa, nn, g


def _shell(s, sh):
    kw = dict(dict(G.SHELLS)[sh])
    kw.setdefault("rout", s.rcut)
    return kw


@pytest.mark.parametrize("sh", ["outer", "switched"])
@pytest.mark.parametrize("name", [st[0] for st in G.STATES])
def test_force_alchemy_matches_reference(ctx, name, sh):
    s, atoms, fa, nn, g = ctx
    lam, intra, ljin = G.state(name)
    al = synth.alchemy_for(s, atoms, intramol=intra, ljintra=ljin)
    r = O.force_alchemy(s, al, lam, nn, fa=fa, **_shell(s, sh))
    assert np.array_equal(np.array([r[k] for k in G.E6] + [r["dwrk"]]), g[f"force_{name}_{sh}_e"])
    assert np.array_equal(r["stress"], g[f"force_{name}_{sh}_stress"])
    assert np.array_equal(digest(r["f"]), g[f"force_{name}_{sh}_f_sha"])


@pytest.mark.parametrize("sh", ["outer", "switched"])
@pytest.mark.parametrize("name", [st[0] for st in G.STATES])
def test_force_alchemy_shift_matches_reference(ctx, name, sh):
    """The shifted-potential form (force_alchemy_shift, SHIFT-POT)."""
    s, atoms, fa, nn, g = ctx
    lam, intra, ljin = G.state(name)
    al = synth.alchemy_for(s, atoms, intramol=intra, ljintra=ljin, shifted=True)
    r = O.force_alchemy(s, al, lam, nn, fa=fa, **_shell(s, sh))
    assert np.array_equal(np.array([r[k] for k in G.E6] + [r["dwrk"]]), g[f"shift_{name}_{sh}_e"])
    assert np.array_equal(r["stress"], g[f"shift_{name}_{sh}_stress"])
    assert np.array_equal(digest(r["f"]), g[f"shift_{name}_{sh}_f_sha"])


@pytest.mark.parametrize("name", [st[0] for st in G.STATES])
def test_corrections_and_pme_match_reference(ctx, name):
    s, atoms, fa, nn, g = ctx
    lam, intra, ljin = G.state(name)
    al = synth.alchemy_for(s, atoms, intramol=intra, ljintra=ljin)
    q = synth.alchemy_charges(s.charge, al.ss_alchemy, lam[4], lam[5])
    q0 = synth.alchemy_charges(s.charge, al.ss_alchemy, lam[6], lam[7])
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    for iz in (1, 0):
        r = O.ferrf_alchemy(s, q, q0, pairs, iz, 0.8)
        assert np.array_equal(np.array([r["slt"], r["slv"], r["dwrk"]]), g[f"ferrf_{name}_{iz}_e"])
        assert np.array_equal(digest(r["f"]), g[f"ferrf_{name}_{iz}_f_sha"])
    assert np.array_equal(np.array(O.cself_alchemy(s, q, q0)), g[f"cself_{name}"])
    # fftpack branch: the FFTs differ from the reference's by rounding only
    p = O.fft_pme_alchemy(s, q, q0, variant="fftpack")
    e = g[f"pme_{name}_e"]
    assert abs(p["eer"] - e[0]) <= 1e-13 * abs(e[0])
    assert abs(p["dwrk"] - e[1]) <= 1e-12 * abs(e[0])
    assert np.abs(p["stress"] - g[f"pme_{name}_stress"]).max() <= 1e-12 * np.abs(g[f"pme_{name}_stress"]).max()
    assert np.abs(p["f"] - g[f"pme_{name}_f"]).max() <= 1e-11 * np.abs(g[f"pme_{name}_f"]).max()


def test_alch_set_tables():
    """alch_set: (soft-core shift, scale) per typeaij = s_i + s_j."""
    f = O.alch_set(0.2, 0.7)
    assert tuple(f[2 + 4]) == (0.0, 1.0) and tuple(f[3 + 4]) == (0.2, 0.8) and tuple(f[-1 + 4]) == (0.7, 1.0 - 0.7)
    assert tuple(f[4 + 4]) == (0.0, 1.0) and tuple(f[-4 + 4]) == (0.0, 1.0) and tuple(f[0 + 4]) == (0.0, 0.0)
    fq = O.alch_set(0.2, 0.7, intramol=True, qq=True)
    assert tuple(fq[4 + 4]) == (0.2 * 0.2, 1.0 - 0.2 * 0.2) and tuple(fq[-4 + 4]) == (0.7 * 0.7, 1.0 - 0.7 * 0.7)
    fl = O.alch_set(0.2, 0.7, intramol=True, ljintra=False)
    assert tuple(fl[-4 + 4]) == (0.7, 1.0 - 0.7)
    fj = O.alch_set(0.2, 0.7, intramol=True, ljintra=True)
    assert tuple(fj[-4 + 4]) == (0.0, 1.0)


@pytest.mark.parametrize("shifted", [False, True])
def test_full_coupling_is_the_plain_pair(ctx, shifted):
    """At lambda = 0 every factor is (0, 1): the soft-core (or shifted) pair is
    force.f's LJ + erfc pair (sigma = rmin_i + rmin_j, 4 sqrt(eps_i eps_j)
    reproduce the Lorentz-Berthelot tables), up to rounding."""
    s, atoms, fa, nn, _ = ctx
    al = synth.alchemy_for(s, atoms, shifted=shifted)
    kw = _shell(s, "switched")
    a = O.force_alchemy(s, al, synth.lambda_state(0.0), nn, fa=fa, **kw)
    b = O.force(s, nn, fa=fa, **kw)
    for k in G.E6:
        assert abs(a[k] - b[k]) <= 1e-13 * max(abs(b[k]), 1e-300), k
    assert np.abs(a["f"] - b["f"]).max() <= 1e-13 * np.abs(b["f"]).max()
    assert np.abs(a["stress"] - b["stress"]).max() <= 1e-13 * np.abs(b["stress"]).max()
    assert a["dwrk"] == 0.0


def test_dwrk_is_the_energy_step(ctx):
    s, atoms, fa, nn, _ = ctx
    al = synth.alchemy_for(s, atoms)
    kw = _shell(s, "outer")

    def u(lrm, lrmq):
        r = O.force_alchemy(s, al, synth.lambda_state(lrm, lrmq=lrmq), nn, fa=fa, **kw)
        return sum(r[k] for k in G.E6)

    r = O.force_alchemy(s, al, synth.lambda_state(0.3, lrm0=0.45, lrmq=0.5, lrmq0=0.2), nn, fa=fa, **kw)
    du = u(0.3, 0.5) - u(0.45, 0.2)
    assert abs(r["dwrk"] - du) <= 1e-12 * abs(u(0.3, 0.5))


@pytest.mark.parametrize("shifted", [False, True])
def test_forces_are_minus_the_energy_gradient(ctx, shifted):
    """Central differences of the total direct energy (soft-core or shifted LJ
    and Coulomb at lambda = 0.5, switched shells) against the forces on
    alchemical and plain atoms; the group switch makes the energy smooth."""
    s, atoms, fa, nn, _ = ctx
    al = synth.alchemy_for(s, atoms, shifted=shifted)
    lam = synth.lambda_state(0.5, lrmq=0.4)
    kw = _shell(s, "switched")
    base = O.force_alchemy(s, al, lam, nn, fa=fa, **kw)
    h = 1e-6
    co = np.asarray(s.co, order="F")
    for a in (0, 5, N := s.natoms - 2):
        for ax in range(3):
            e = []
            for sgn in (1, -1):
                d = np.zeros(3)
                d[ax] = sgn * h
                dp = np.linalg.solve(co, d)  # reduced displacement
                xs = [s.xp0.copy(), s.yp0.copy(), s.zp0.copy()]
                for k in range(3):
                    xs[k][a] += dp[k]
                t = s.with_coords(*xs)
                r = O.force_alchemy(t, al, lam, nn, fa=O.ForceArgs(t), **kw)
                e.append(sum(r[k] for k in G.E6))
            fd = -(e[0] - e[1]) / (2 * h)
            assert abs(fd - base["f"][ax, a]) <= 1e-6 * np.abs(base["f"]).max() + 1e-5 * abs(fd), (a, ax, N)


def test_decoupled_ligand_leaves_the_solvent_alone(ctx):
    """At lambda = 1 (removed atoms fully off, intramol off) every pair of a
    ligand atom with another atom is scaled by 0: the solvent-solvent energies
    are unchanged, and the forces on the other atoms are those of the fully
    coupled system whose ligand has zero charges and zero epsilon."""
    s, atoms, fa, nn, _ = ctx
    al = synth.alchemy_for(s, atoms)
    kw = _shell(s, "outer")
    r = O.force_alchemy(s, al, synth.lambda_state(1.0), nn, fa=fa, **kw)
    w = O.force_alchemy(s, al, synth.lambda_state(0.0), nn, fa=fa, **kw)
    assert r["uconf_slv"] == w["uconf_slv"] and r["ucoul_slv"] == w["ucoul_slv"]
    q = s.charge.copy()
    q[atoms] = 0.0
    epsm = al.epsm.copy()
    epsm[atoms] = 0.0
    z = O.force_alchemy(dataclasses.replace(s, charge=q), dataclasses.replace(al, epsm=epsm),
                        synth.lambda_state(0.0), nn, fa=fa, **kw)
    others = np.setdiff1d(np.arange(s.natoms), atoms)
    assert np.abs(r["f"][:, others] - z["f"][:, others]).max() <= 1e-13 * np.abs(z["f"]).max()
