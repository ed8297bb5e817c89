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
