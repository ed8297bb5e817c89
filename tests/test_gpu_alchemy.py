"""Alchemical force path on the GPU against the oracle (FP64 bar: forces
relative RMS <= 1e-9, energies <= 1e-10; dwrk, a difference of energies,
within 1e-10 of the energies' scale; mixed precision: forces <= 5e-6).

* orac_hip_force_alchemy (force_alchemy.f's interface) on the lambda states
  of tests/golden/make_alchemy.py, single and switched shells;
* a resident alchemical system: soft-core pairs + reciprocal sum on the
  lambda charges + ferrf corrections, and dwrk from the lambda0 charges
  (energy-only grid), against force_alchemy + fft_pme_alchemy +
  ferrf_alchemy + cself_alchemy;
* lambda-windows batched: each window the same as alone.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from conftest import rel, relrms

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_alchemy as G  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def alch_sys():
    from orac_hip import synth
    s, atoms = G.system()
    fa = O.ForceArgs(s)
    return s, atoms, fa, G.outer_list(s, fa), synth


def _gpu_force_alchemy(ctx, s, al, lam, nnlpp, fa, *, rinn, rout, rtolinn, rtolout, rneigh):
    n = s.natoms
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    st = np.zeros((3, 3), order="F")
    nn1 = np.zeros(len(nnlpp) + s.ngrp + 16, np.int32)
    r = ctx.force_alchemy(s.ss_index, s.xp0, s.yp0, s.zp0, fa.xpg, fa.ypg, fa.zpg, s.charge, s.nbtype, s.type_table,
                          s.ntypes, n, s.atomg, fa.xpcm, fa.ypcm, fa.zpcm, None, s.atomp, s.co, s.ecc12, s.ecc6,
                          s.alpha, fa.mapnl, s.ngrp, s.grppt, fpx, fpy, fpz, st, nnlpp, nn1, rneigh, rinn, rout,
                          rtolinn, rtolout, s.pmass, 1, s.mass, 1.0e20, False, True, alchemy=al, lam=lam, dwrk=0.0)
    r.update(f=np.stack([fpx, fpy, fpz]), stress=st)
    return r


@pytest.mark.parametrize("shifted", [False, True])
@pytest.mark.parametrize("sh", ["outer", "switched"])
@pytest.mark.parametrize("name", [st[0] for st in G.STATES])
def test_force_alchemy_entry(gpu_ctx, alch_sys, name, sh, shifted):
    """orac_hip_force_alchemy (soft core) and orac_hip_force_alchemy_shift."""
    s, atoms, fa, nn, synth = alch_sys
    lam, intra, ljin = G.state(name)
    al = synth.alchemy_for(s, atoms, intramol=intra, ljintra=ljin, shifted=shifted)
    kw = dict(dict(G.SHELLS)[sh])
    kw.setdefault("rout", s.rcut)
    g = _gpu_force_alchemy(gpu_ctx, s, al, lam, nn, fa, **kw)
    o = O.force_alchemy(s, al, lam, nn, fa=fa, **kw)
    err = relrms(g["f"], o["f"])
    print(name, sh, "f relrms", err, "dwrk", g["dwrk"], o["dwrk"])
    assert err < 1e-9
    scale = sum(abs(o[k]) for k in G.E6)
    for k in G.E6:
        assert abs(g[k] - o[k]) <= 1e-10 * max(abs(o[k]), 1e-3 * scale), k
    assert abs(g["dwrk"] - o["dwrk"]) <= 1e-10 * scale
    assert np.abs(g["stress"] - o["stress"]).max() <= 1e-10 * np.abs(o["stress"]).max()


def oracle_alchemy_total(s, al, lam, cp, cf, synth):
    """The alchemical force step of a resident system on the oracle."""
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    nn = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0,
                 fa=fa)["nnlpp1"]
    d = O.force_alchemy(s, al, lam, nn, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0)
    q = synth.alchemy_charges(s.charge, al.ss_alchemy, lam[4], lam[5])
    q0 = synth.alchemy_charges(s.charge, al.ss_alchemy, lam[6], lam[7])
    p = O.fft_pme_alchemy(s, q, q0, variant="fftw", nthreads=4)
    f = d["f"] + p["f"]
    cdw = 0.0
    for fud in np.unique(cf):
        sel = cf == fud
        c = O.ferrf_alchemy(s, q, q0, cp[:, sel], 1, float(fud))
        f = f + c["f"]
        cdw += c["dwrk"]
    if al.listqq is not None:
        # ferrf_alchemy_add.f on the intra-alchemical pairs (mtsmd.f:3034-3044)
        lams = synth.alchemy_lambdas(al.ss_alchemy, lam)
        c = O.ferrf_alchemy_add(s, lams[1], lams[3], al.listqq, intramol=al.intramol)
        f = f + c["f"]
        cdw += c["dwrk"]
    self_slt, sdw = O.cself_alchemy(s, q, q0)
    return dict(f=f, d=d, p=p, corr_dw=cdw, self_slt=self_slt, self_dw=sdw,
                dwrk=d["dwrk"] + p["dwrk"] + cdw + sdw)


@pytest.mark.parametrize("precision,shifted,intra", [(0, False, False), (1, False, False), (0, True, False),
                                                     (1, True, False), (0, False, True)])
def test_resident_alchemy_matches_oracle(gpu_ctx, alch_sys, precision, shifted, intra):
    """The resident alchemical step, with ferrf_alchemy_add on the 12 removed
    atoms' intra-alchemical pairs (bonds and 1-3 pairs excluded)."""
    from orac_hip.api import ResidentSystem
    s, atoms, _, _, synth = alch_sys
    cp = np.concatenate([s.bonds, s.bends], axis=1)
    cf = np.ones(cp.shape[1])
    al = synth.alchemy_for(s, atoms, shifted=shifted, intramol=intra)
    al.listqq = synth.alchemy_qq_list([], atoms + 1, (s.bonds, s.bends))
    assert al.listqq.shape[1] > 30
    rs = ResidentSystem(gpu_ctx, s, precision=precision, ncell=2, corr_pairs=cp, corr_fudge=cf)
    rs.set_alchemy(al)
    for lrm, lrm0, lrmq, lrmq0 in [(0.35, 0.4, 0.6, 0.55), (0.0, 0.05, 0.0, 0.05), (1.0, 0.9, 1.0, 0.9)]:
        lam = synth.lambda_state(lrm, lrm0=lrm0, lrmq=lrmq, lrmq0=lrmq0)
        rs.set_lambda(lam)
        rs.eval(rebuild_list=True)
        r = rs.results()
        a = rs.alchemy_results()
        o = oracle_alchemy_total(s, al, lam, cp, cf, synth)
        err = relrms(r["f"], o["f"])
        scale = abs(o["p"]["eer"]) + sum(abs(o["d"][k]) for k in G.E6)
        print(precision, lrm, "f relrms", err, "dwrk", a["dwrk"], o["dwrk"], a)
        if precision == 0:
            assert err < 1e-9
            assert rel(r["eer"], o["p"]["eer"]) < 1e-10
            assert abs(a["eer0"] - o["p"]["eer0"]) <= 1e-10 * abs(o["p"]["eer0"])
            assert abs(a["direct"] - o["d"]["dwrk"]) <= 1e-10 * scale
            assert abs(a["recip"] - o["p"]["dwrk"]) <= 1e-10 * scale
            assert abs(a["corr"] - o["corr_dw"]) <= 1e-10 * scale
            assert abs(a["dwrk"] - o["dwrk"]) <= 1e-10 * scale
            assert rel(r["self_slt"], o["self_slt"]) < 1e-12
        else:
            assert err < 5e-6
            assert abs(a["dwrk"] - o["dwrk"]) <= 1e-5 * scale
    rs.close()


def test_windows_batched_equal_alone(gpu_ctx):
    """Four lambda windows of one small system in one batch: forces and dwrk
    of each window the same as the window evaluated alone."""
    from orac_hip import synth
    from orac_hip.api import Batch, ResidentSystem
    s, atoms = G.system()
    al = synth.alchemy_for(s, atoms)
    lams = [synth.lambda_state(l, lrm0=min(1.0, l + 0.1)) for l in (0.0, 0.3, 0.6, 0.9)]
    alone = []
    for lam in lams:
        rs = ResidentSystem(gpu_ctx, s, precision=0, ncell=2)
        rs.set_alchemy(al, lam)
        rs.eval(rebuild_list=True)
        alone.append((rs.results()["f"], rs.alchemy_results()["dwrk"]))
        rs.close()
    rss = []
    for lam in lams:
        rs = ResidentSystem(gpu_ctx, s, precision=0, ncell=2)
        rs.set_alchemy(al, lam)
        rss.append(rs)
    b = Batch(gpu_ctx, rss)
    b.eval(rebuild_list=True)
    for rs, (f, dw) in zip(rss, alone):
        assert np.array_equal(rs.results()["f"], f)
        assert abs(rs.alchemy_results()["dwrk"] - dw) <= 1e-12 * max(1.0, abs(dw))
    b.close()
    for rs in rss:
        rs.close()


def test_plain_reciprocal_and_corrections(gpu_ctx, alch_sys):
    """orac_hip_fft_pme_alchemy, _ferrf_alchemy and _cself_alchemy (the routines
    mts_furier_alchemy.f calls) against the oracle."""
    s, atoms, _, _, synth = alch_sys
    al = synth.alchemy_for(s, atoms)
    q = synth.alchemy_charges(s.charge, al.ss_alchemy, 1.0, 0.35)
    q0 = synth.alchemy_charges(s.charge, al.ss_alchemy, 1.0, 0.4)
    n = s.natoms
    gpu_ctx.fft_pme_init(n, *s.nfft, s.order)
    cm = s.molecule_coms()
    fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
    st = np.zeros((3, 3), order="F")
    eer, dw = gpu_ctx.fft_pme_alchemy(n, s.xp0, s.yp0, s.zp0, cm[0], cm[1], cm[2], q, q0, s.co, s.oc, s.volume,
                                      s.alpha, s.order, *s.nfft, fpx, fpy, fpz, st, s.atomp, s.grppt, True, 1e9)
    o = O.fft_pme_alchemy(s, q, q0, variant="fftw", nthreads=4)
    assert rel(eer, o["eer"]) < 1e-10
    assert abs(dw - o["dwrk"]) <= 1e-10 * abs(o["eer"])
    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-9
    assert np.abs(st - o["stress"]).max() <= 1e-10 * np.abs(o["stress"]).max()
    pairs = np.asfortranarray(np.concatenate([s.bonds, s.bends], axis=1))
    x, y, z = s.cartesian()
    for iz in (1, 0):
        fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
        a, b, d = gpu_ctx.ferrf_alchemy(s.ss_index, s.alpha, q, q0, 0.8, x, y, z, iz, pairs, pairs.shape[1], fpx,
                                        fpy, fpz)
        r = O.ferrf_alchemy(s, q, q0, pairs, iz, 0.8)
        assert rel(a, r["slt"]) < 1e-12 and b == r["slv"]
        assert abs(d - r["dwrk"]) <= 1e-12 * abs(r["slt"])
        if iz == 1:
            assert relrms(np.stack([fpx, fpy, fpz]), r["f"]) < 1e-12
    a, d = gpu_ctx.cself_alchemy(s.ss_index, n, s.alpha, 1e9, q, q0)
    ra, rd = O.cself_alchemy(s, q, q0)
    assert rel(a, ra) < 1e-13 and abs(d - rd) <= 1e-13 * abs(ra)


def test_plain_intra_alchemical_terms(gpu_ctx):
    """orac_hip_ferrf_alchemy_add (plain and ERF_CORR, intramol on and off),
    orac_hip_fnb14_alchemy (Ewald and plain) and orac_hip_fnbgrp_alchemy
    against the oracle on the added/removed solute of
    tests/golden/make_alchemy_intra.py, for every lambda state there
    (forces relative RMS <= 1e-12, energies and dwrk <= 1e-12 of their scale)."""
    import make_alchemy_intra as GI
    s, ss = GI.system()
    n = s.natoms
    lq = GI.qq_list(s)
    p14, t14 = GI.pairs14(s)
    x, y, z = s.cartesian()
    nb = 400
    tab, _ = gpu_ctx.int_corr_erf_spline(0.5, 12.0, nb, s.alpha, 2.0)   # the ERF_CORR table (host builder)
    delew, rlew = (12.0 - 0.5) / nb, 0.5
    for st in GI.STATES:
        lam, lamq, lam0, lamq0 = GI.lams(ss, st)
        for intra in (False, True):
            for ec in (False, True):
                fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
                a, b, dw = gpu_ctx.ferrf_alchemy_add(s.ss_index, s.alpha, s.charge, lamq, lamq0, x, y, z, lq,
                                                     lq.shape[1], fpx, fpy, fpz, erf_corr=ec,
                                                     erf_arr_corr=tab if ec else None, delew=delew, rlew=rlew,
                                                     intramol=intra)
                o = O.ferrf_alchemy_add(s, lamq, lamq0, lq, intramol=intra, erf=(tab, delew, rlew) if ec else None)
                # dwrk is its own scale where intramol leaves both energies exactly zero
                scale = abs(o["slt"]) + abs(o["slv"]) + abs(o["dwrk"]) + 1e-30
                print(st[0], intra, ec, a, o["slt"], dw, o["dwrk"])
                assert abs(a - o["slt"]) <= 1e-12 * scale and abs(b - o["slv"]) <= 1e-12 * scale
                assert abs(dw - o["dwrk"]) <= 1e-12 * scale
                if np.abs(o["f"]).max() > 0:
                    assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
        for ew in (True, False):
            fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
            e = gpu_ctx.fnb14_alchemy(s.ss_index, x, y, z, s.charge, n, s.ecc12, s.ecc6, True, ew, s.alpha, p14,
                                      p14.shape[1], t14, *GI.FNB_FUDGE, fpx, fpy, fpz, lam, lamq, lam0, lamq0)
            o = O.fnb14_alchemy(s, p14, t14, *GI.FNB_FUDGE, (lam, lamq, lam0, lamq0), ewald=ew)
            scale = sum(abs(o[k]) for k in O.FNB_E)
            for v, k in zip(e[:4], O.FNB_E):
                assert abs(v - o[k]) <= 1e-12 * scale, k
            assert abs(e[4] - o["dwrk"]) <= 1e-12 * scale
            assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
            fpx, fpy, fpz = np.zeros(n), np.zeros(n), np.zeros(n)
            e = gpu_ctx.fnbgrp_alchemy(s.ss_index, x, y, z, s.charge, s.nbtype, s.ecc12, s.ecc6, ew, s.alpha, p14,
                                       p14.shape[1], fpx, fpy, fpz, lamq)
            o = O.fnbgrp_alchemy(s, p14, s.nbtype, ew, lamq)
            scale = sum(abs(o[k]) for k in O.FNB_E)
            for v, k in zip(e, O.FNB_E):
                assert abs(v - o[k]) <= 1e-12 * scale, k
            assert relrms(np.stack([fpx, fpy, fpz]), o["f"]) < 1e-12
