"""The Fortran drop-in: a Fortran caller (tests/fortran/drive_shim.f90) calls
fft_pme_init, fft_pme, lc_index, lc_list, force, ferrf and cself with the
reference's implicit-interface argument lists, linked against the shim
(orac-porting_amd/fortran/orac_hip_shim.f90) and the HIP library."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, rel, relrms

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "fortran", "drive_shim")


def write_input(path, s, ncx):
    fa = O.ForceArgs(s)
    x, y, z = s.cartesian()
    corr = np.concatenate([s.bonds, s.bends], axis=1) if s.bonds is not None else np.zeros((2, 0), np.int32)
    cons = s.constraints
    mapnl = s.mapnl()
    with open(path, "wb") as f:
        np.array([s.natoms, s.ngrp, s.nprot, s.ntypes, *s.nfft, s.order, len(mapnl), corr.shape[1], cons.shape[1],
                  ncx], np.int32).tofile(f)
        np.array([s.alpha, s.volume, s.rcut, s.rtol, s.rnei], np.float64).tofile(f)
        for a in (s.co, s.oc):
            np.asfortranarray(a, dtype=np.float64).ravel(order="F").tofile(f)
        for a in (s.xp0, s.yp0, s.zp0, s.charge, s.mass, s.pmass, fa.xpg, fa.ypg, fa.zpg, fa.xpcm, fa.ypcm, fa.zpcm,
                  s.ecc12, s.ecc6, x, y, z):
            np.asarray(a, np.float64).tofile(f)
        for a in (s.nbtype, s.atomg, s.atomp, s.ss_index):
            np.asarray(a, np.int32).tofile(f)
        np.asarray(s.type_table, np.int32).ravel(order="F").tofile(f)
        np.asarray(s.grppt, np.int32).ravel(order="F").tofile(f)
        np.asarray(mapnl, np.int32).tofile(f)
        if corr.shape[1]:
            np.asarray(corr, np.int32).ravel(order="F").tofile(f)
        np.asarray(cons, np.int32).ravel(order="F").tofile(f)
    return corr, cons


def read_output(path, n):
    b = open(path, "rb").read()
    d = np.frombuffer(b[: (10 + 3 * n + 6 + 9 + 3 * n + 6) * 8], np.float64)
    o = {}
    o["eer"] = d[0]
    o["stressc"] = d[1:10].reshape(3, 3, order="F")
    o["fpme"] = d[10:10 + 3 * n].reshape(3, n)
    k = 10 + 3 * n
    o["e"] = d[k:k + 6]
    o["stress"] = d[k + 6:k + 15].reshape(3, 3, order="F")
    o["fdir"] = d[k + 15:k + 15 + 3 * n].reshape(3, n)
    k = k + 15 + 3 * n
    o["corr"] = d[k:k + 6]
    return o


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (flang absent at build time)")
def test_fortran_caller_through_shim(tmp_path, small_solute):
    s = small_solute
    ncx = int(s.L // 3.0)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    corr, cons = write_input(inp, s, ncx)
    r = subprocess.run([EXE, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    g = read_output(out, s.natoms)
    p = O.fft_pme(s, variant="fftw")
    assert rel(g["eer"], p["eer"]) < 1e-10
    assert relrms(g["fpme"], p["f"]) < 1e-9
    assert np.abs(g["stressc"] - p["stress"]).max() < 1e-10 * np.abs(p["stress"]).max()
    rc = s.rcut + s.rtol + s.rnei
    fa = O.ForceArgs(s)
    st = O.lc_index(ncx, ncx, ncx, rc, s.co)
    nnlpp = O.lc_list(ncx, ncx, ncx, st, rc, s.co, fa.xpg, fa.ypg, fa.zpg)
    d = O.force(s, nnlpp, fa=fa, rinn=0.0, rout=s.rcut, rtolinn=0.0, rtolout=s.rtol, rneigh=0.0)
    c1 = O.ferrf(s, corr, 1, 1.0, fp=d["f"])
    names = ["uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss"]
    for v, k in zip(g["e"], names):
        assert (v == 0.0 and d[k] == 0.0) or rel(v, d[k]) < 1e-10, k
    assert relrms(g["fdir"], c1["f"]) < 1e-9
    c0 = O.ferrf(s, cons, 0, 1.0)
    cs = O.cself(s)
    assert rel(g["corr"][0], c1["slt"]) < 1e-10
    assert rel(g["corr"][3], c0["slv"]) < 1e-10
    assert rel(g["corr"][5], cs[1]) < 1e-12


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (flang absent at build time)")
def test_fortran_caller_error_reaches_xerror(tmp_path, water216):
    """A stencil that would visit a cell twice is fatal in the reference
    (lc_index.f:106-122); through the shim it reaches the caller's xerror."""
    s = water216
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(inp, s, 6)
    r = subprocess.run([EXE, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "xerror" in r.stderr and "lc_index" in r.stderr


INTRA = os.path.join(ROOT, "tests", "fortran", "drive_intra")


@pytest.mark.skipif(not os.path.exists(INTRA), reason="Fortran driver not built (flang absent at build time)")
def test_fortran_intra_and_alchemy_through_shim(tmp_path):
    """A Fortran caller (tests/fortran/drive_intra.f90) calls fpbond, fpbend,
    fptors, fitors, fnb14 (itask 1 and 2), fnbgrp, ferrf_alchemy_add,
    fnb14_alchemy, fnbgrp_alchemy, ferrf_alchemy and cself_alchemy through
    the shim on the FKBP12 topology (atoms 1-12 removed and 13-24 added
    alchemically), against the oracle (forces relative RMS <= 1e-12,
    energies <= 1e-12 of their scale)."""
    from orac_hip import formats, synth
    s, _ = formats.fkbp12()
    g = np.load(formats.FKBP12_PATH)
    n = s.natoms
    x, y, z = s.cartesian()
    i32 = lambda a: np.asarray(a, np.int32)  # noqa: E731
    lb, bp = i32(g["lbnd"]), g["bond_pots"]
    lbg, bep = i32(g["lbndg"]), g["bend_pots"]
    lt, pt = i32(g["ltor"]), g["potto"]
    li, ip = i32(g["litr"]), g["itor_pots"]
    i14, t14 = i32(g["int14"]), i32(g["type14"])
    ingrp = i32(g["fnb_intra"])
    ss_al = np.ones(n, np.int32)
    ss_al[:12], ss_al[12:24] = -2, 2
    lamv = np.array([0.3, 0.6, 0.35, 0.55, 0.4, 0.7, 0.45, 0.65])
    lams = synth.alchemy_lambdas(ss_al, lamv)
    qq = synth.alchemy_qq_list(np.arange(13, 25), np.arange(1, 13), (s.bonds, s.bends, i14))
    qa = synth.alchemy_charges(s.charge, ss_al, lamv[4], lamv[5])
    qa0 = synth.alchemy_charges(s.charge, ss_al, lamv[6], lamv[7])
    lbc = i32(np.concatenate([s.bonds, s.bends], axis=1))
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(inp, "wb") as f:
        np.array([n, lb.shape[1], lbg.shape[1], lt.shape[1], li.shape[1], i14.shape[1], qq.shape[1], ingrp.shape[1],
                  len(s.ecc12), lbc.shape[1]], np.int32).tofile(f)
        np.array([s.alpha, 0.83333, 0.5], np.float64).tofile(f)
        i32(s.ss_index).tofile(f)
        for a in (x, y, z, s.charge):
            np.asarray(a, np.float64).tofile(f)
        i32(s.nbtype).tofile(f)
        for a in (s.ecc12, s.ecc6):
            np.asarray(a, np.float64).tofile(f)
        for a, t in ((lb, np.int32), (bp, np.float64), (lbg, np.int32), (bep, np.float64), (lt, np.int32),
                     (pt, np.float64), (li, np.int32), (ip, np.float64), (i14, np.int32), (t14, np.int32),
                     (ingrp, np.int32), (qq, np.int32), (lbc, np.int32)):
            np.asarray(a, t).ravel(order="F").tofile(f)
        for a in (*lams, qa, qa0):
            np.asarray(a, np.float64).tofile(f)
    r = subprocess.run([INTRA, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    rec = np.fromfile(out, np.float64).reshape(12, 5 + 3 * n)
    E = [rr[:5] for rr in rec]
    F = [rr[5:].reshape(3, n) for rr in rec]

    def check(k, o, names, extra=()):
        vals = [o[nm] for nm in names] + list(extra)
        scale = sum(abs(v) for v in vals) + 1e-300
        for got, want in zip(E[k], vals):
            assert abs(got - want) <= 1e-12 * scale, (k, got, want)
        if "f" in o and np.abs(o["f"]).max() > 0:
            assert relrms(F[k], o["f"]) < 1e-12, k

    check(0, O.fpbond(s, lb, bp[0], bp[1]), ["ubond_slt", "ubond_slv"])
    check(1, O.fpbend(s, lbg, list(bep)), ["ubend_slt", "ubend_slv"])
    check(2, O.fptors(s, lt, pt[:, 1], pt[:, 0]), ["utors_slt", "utors_slv"])
    check(3, O.fitors(s, li, *ip), ["utors_slt", "utors_slv"])
    fnb = ["uconf_slt", "uconf_slv", "ucoul_slt", "ucoul_slv"]
    check(4, O.fnb14(s, i14, t14, 0.83333, 0.5), fnb)
    check(5, O.fnbgrp(s, ingrp, s.nbtype, True), fnb)
    assert np.all(E[6] == 0.0) and np.all(F[6] == 0.0)          # itask 2: another thread's share is empty
    o = O.ferrf_alchemy_add(s, lams[1], lams[3], qq)
    check(7, o, ["slt", "slv", "dwrk"])
    o = O.fnb14_alchemy(s, i14, t14, 0.83333, 0.5, lams)
    check(8, o, fnb + ["dwrk"])
    check(9, O.fnbgrp_alchemy(s, ingrp, s.nbtype, True, lams[1]), fnb)
    check(10, O.ferrf_alchemy(s, qa, qa0, lbc, 1, 0.8), ["slt", "slv", "dwrk"])
    a, d = O.cself_alchemy(s, qa, qa0)
    assert abs(E[11][0] - a) <= 1e-13 * abs(a) and abs(E[11][1] - d) <= 1e-13 * abs(a)
