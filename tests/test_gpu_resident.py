"""Device-resident full force evaluation against the oracle's composition of
force (single shell) + fft_pme (FFTW branch) + ferrf (per-step corrections).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel, relrms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solvated24k():
    from orac_hip import synth
    return synth.config("solvated24k")


def oracle_total(s, corr_pairs, corr_fudge, nthreads=4):
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    if s.L > 3 * rc:
        nc = int(s.L // 3.0)
        st = O.lc_index(nc, nc, nc, rc, s.co)
        nnlpp = O.lc_list(nc, nc, nc, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
    else:
        nnlpp = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0,
                        fa=fa)["nnlpp1"]
    d = O.force(s, nnlpp, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0, nthreads=nthreads)
    p = O.fft_pme(s, variant="fftw", nthreads=nthreads)
    f = d["f"] + p["f"]
    cs = cv = 0.0
    if corr_pairs.shape[1]:
        # group by fudge value so each ferrf call has one fudge
        for fud in np.unique(corr_fudge):
            sel = corr_fudge == fud
            c = O.ferrf(s, corr_pairs[:, sel], 1, float(fud))
            f = f + c["f"]
            cs += c["slt"]
            cv += c["slv"]
    stress = d["stress"] + p["stress"]
    return dict(f=f, eer=p["eer"], d=d, corr_slt=cs, corr_slv=cv, stress=stress)


def _corr(s):
    if s.bonds is None:
        return np.zeros((2, 0), np.int32), np.zeros(0)
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    return pairs, np.ones(pairs.shape[1])


@pytest.mark.parametrize("sysname", ["water216", "small_solute", "solvated24k"])
def test_resident_fp64_matches_oracle(gpu_ctx, sysname, request):
    from orac_hip.api import ResidentSystem
    s = request.getfixturevalue(sysname)
    cp, cf = _corr(s)
    rs = ResidentSystem(gpu_ctx, s, precision=0, corr_pairs=cp, corr_fudge=cf)
    rs.eval(rebuild_list=True)
    g = rs.results()
    o = oracle_total(s, cp, cf)
    print(sysname, "f relrms", relrms(g["f"], o["f"]), "eer", rel(g["eer"], o["eer"]))
    assert relrms(g["f"], o["f"]) < 1e-9
    assert rel(g["eer"], o["eer"]) < 1e-10
    for k in ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"):
        if o["d"][k] != 0.0:
            assert rel(g[k], o["d"][k]) < 1e-10, k
    if cp.shape[1]:
        assert rel(g["corr_slt"], o["corr_slt"]) < 1e-10
    assert np.abs(g["stress"] - o["stress"]).max() < 1e-10 * np.abs(o["stress"]).max()
    cs = O.cself(s)
    assert rel(g["self_slv"], cs[1]) < 1e-12
    # second evaluation, no rebuild: bit-identical
    rs.eval(rebuild_list=False)
    g2 = rs.results()
    assert np.array_equal(g["f"], g2["f"]) and g["eer"] == g2["eer"]
    rs.close()


def test_resident_mixed_precision_accuracy(gpu_ctx, solvated24k):
    from orac_hip.api import ResidentSystem
    s = solvated24k
    cp, cf = _corr(s)
    rs = ResidentSystem(gpu_ctx, s, precision=1, corr_pairs=cp, corr_fudge=cf)
    rs.eval()
    g = rs.results()
    o = oracle_total(s, cp, cf)
    err = relrms(g["f"], o["f"])
    print("mixed f relrms", err, "eer", rel(g["eer"], o["eer"]))
    assert err < 5e-6
    assert rel(g["eer"], o["eer"]) < 1e-5
    rs.close()


def test_resident_moves_and_rebuilds(gpu_ctx, small_solute):
    """Coordinates change between steps: the rebuilt list and forces follow the oracle."""
    from orac_hip.api import ResidentSystem
    s = small_solute
    cp, cf = _corr(s)
    rs = ResidentSystem(gpu_ctx, s, precision=0, corr_pairs=cp, corr_fudge=cf)
    rng = np.random.default_rng(3)
    for step in range(3):
        d = rng.normal(scale=0.02 / (s.L / 2), size=(3, s.natoms))
        s = s.with_coords(s.xp0 + d[0], s.yp0 + d[1], s.zp0 + d[2])
        rs.set_coords(s.xp0, s.yp0, s.zp0)
        rs.eval(rebuild_list=True)
        g = rs.results()
        o = oracle_total(s, cp, cf)
        assert relrms(g["f"], o["f"]) < 1e-9, step
    rs.close()


@pytest.mark.parametrize("split", [1, 2, 4])
def test_pair_split_keeps_forces(gpu_ctx, solvated24k, split, monkeypatch):
    """S waves per owner group split its neighbour list in contiguous parts;
    the owner's force is summed over the parts in a fixed order, so results
    agree with the oracle for every S and differ between S only by rounding."""
    from orac_hip.api import ResidentSystem
    s = solvated24k
    cp, cf = _corr(s)
    monkeypatch.setenv("ORAC_HIP_PAIR_SPLIT", str(split))
    rs = ResidentSystem(gpu_ctx, s, precision=0, corr_pairs=cp, corr_fudge=cf)
    rs.eval(rebuild_list=True)
    g = rs.results()
    rs.eval(rebuild_list=True)
    assert np.array_equal(g["f"], rs.results()["f"])  # deterministic for a given S
    rs.close()
    o = oracle_total(s, cp, cf)
    assert relrms(g["f"], o["f"]) < 1e-9
    for k in ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"):
        if o["d"][k] != 0.0:
            assert rel(g[k], o["d"][k]) < 1e-10, k


def test_packed_lj_table_path(gpu_ctx, small_solute):
    """More LJ types than the dense table takes (2,049+): the pair kernel reads
    the reference's packed type_table and ecc12/ecc6.  The system's own types
    are moved to the end of a 2,100-type table with random other entries."""
    import dataclasses
    from orac_hip import synth
    from orac_hip.api import ResidentSystem
    s = small_solute
    nt0, nt = s.ntypes, 2100
    rng = np.random.default_rng(5)
    sig, eps = rng.uniform(2.0, 3.5, nt), rng.uniform(0.01, 0.3, nt)
    tt, e12, e6 = synth._lj_tables(sig, eps)
    # types t (1-based) of the system become nt - nt0 + t; copy their pair parameters
    off = nt - nt0
    for i in range(nt0):
        for j in range(nt0):
            src = s.type_table[i, j] - 1
            dst = tt[off + i, off + j] - 1
            e12[dst], e6[dst] = s.ecc12[src], s.ecc6[src]
    s2 = dataclasses.replace(s, ntypes=nt, type_table=tt, ecc12=e12, ecc6=e6,
                             nbtype=(s.nbtype + off).astype(np.int32))
    cp, cf = _corr(s2)
    rs = ResidentSystem(gpu_ctx, s2, precision=0, corr_pairs=cp, corr_fudge=cf)
    rs.eval(rebuild_list=True)
    g = rs.results()
    rs.close()
    o = oracle_total(s2, cp, cf)
    assert relrms(g["f"], o["f"]) < 1e-9
    for k in ("uconf_slt", "uconf_slv", "uconf_ss"):
        if o["d"][k] != 0.0:
            assert rel(g[k], o["d"][k]) < 1e-10, k
    # and the same physics as with the system's own small table
    rs = ResidentSystem(gpu_ctx, s, precision=0, corr_pairs=cp, corr_fudge=cf)
    rs.eval(rebuild_list=True)
    assert relrms(rs.results()["f"], g["f"]) < 1e-12
    rs.close()


# ---------------------------------------------------------------- owner-atom loop form (opt-in)

@pytest.mark.parametrize("sysname", ["water216", "solvated24k"])
def test_owner_form_half_list_matches_oracle(gpu_ctx, sysname, request, monkeypatch):
    """ORAC_HIP_PAIR_KERNEL=own: resident systems of groups of at most three
    atoms take the half-list layout and the owner-atom loop kernel; FP64
    forces, energies and virial against the oracle at the plain path's
    tolerances, and bit-identical on a second evaluation."""
    from orac_hip.api import ResidentSystem
    monkeypatch.setenv("ORAC_HIP_PAIR_KERNEL", "own")
    s = request.getfixturevalue(sysname)
    cp, cf = _corr(s)
    rs = ResidentSystem(gpu_ctx, s, precision=0, corr_pairs=cp, corr_fudge=cf)
    rs.eval(rebuild_list=True)
    g = rs.results()
    rs.eval(rebuild_list=True)
    assert np.array_equal(g["f"], rs.results()["f"])
    rs.close()
    o = oracle_total(s, cp, cf)
    assert relrms(g["f"], o["f"]) < 1e-9
    for k in ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"):
        if o["d"][k] != 0.0:
            assert rel(g[k], o["d"][k]) < 1e-10, k
    assert np.abs(g["stress"] - o["stress"]).max() <= 1e-10 * np.abs(o["stress"]).max()


def test_owner_form_mixed_batch_equals_alone(gpu_ctx, monkeypatch):
    """The half-list layout in mixed precision: a batch of three systems gives
    each one its forces alone, bit for bit."""
    from orac_hip import synth
    from orac_hip.api import Batch, ResidentSystem
    monkeypatch.setenv("ORAC_HIP_PAIR_KERNEL", "own")
    systems = [ResidentSystem(gpu_ctx, s, precision=1) for s in
               (synth.config("water216"), synth.small_solvated(rcut=7.0),
                synth.water_box(seed=77, L=24.0, nside=8, nfft=(24, 30, 25), rcut=9.0))]
    alone = []
    for rs in systems:
        rs.eval(rebuild_list=True)
        alone.append(rs.results()["f"])
    b = Batch(gpu_ctx, systems)
    b.eval(rebuild_list=True)
    for r, a in zip(b.results(), alone):
        assert np.array_equal(r["f"], a)
    b.close()
    for x in systems:
        x.close()
