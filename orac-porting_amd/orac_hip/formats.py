"""Readers of ORAC's on-disk formats: the ASCII topology-parameter dump
(``tpgprm``, written and read record by record as ``readtp.f:45-228`` lists
them) and the ASCII restart file (``dumprs.f:118-147`` / ``readrs.f:216-233``),
the writer of that restart file, and the assembly of a ``synth.System`` from
the two.

Both files are Fortran list-directed text: every READ starts on a new line,
takes as many values as it needs from that line and the following ones, and
drops the rest of its last line; ``r*v`` stands for r copies of v.  The
reader follows that rule exactly, so records of any length parse the same.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import synth


class _ListDirected:
    def __init__(self, path):
        with open(path) as f:
            self.lines = f.read().split("\n")
        self.i = 0

    def values(self, n):
        """The next READ of n values (a READ of nothing still consumes one line)."""
        if n == 0:
            self.i += 1
            return []
        out = []
        while len(out) < n:
            if self.i >= len(self.lines):
                raise ValueError("unexpected end of file")
            for t in self.lines[self.i].split():
                if "*" in t and not t.startswith("'"):
                    r, v = t.split("*", 1)
                    out.extend([v] * int(r))
                else:
                    out.append(t)
            self.i += 1
        return out[:n]

    def ints(self, n):
        return np.array([int(v) for v in self.values(n)], dtype=np.int64)

    def reals(self, n):
        return np.array([float(v.replace("D", "E").replace("d", "e")) for v in self.values(n)])

    def counted_lists(self, n):
        """One READ of n lists, each a count followed by that many integers (concta)."""
        out, buf = [], []
        while len(out) < n:
            if not buf:
                buf = self.lines[self.i].split()
                self.i += 1
                continue
            c = int(buf.pop(0))
            vals = []
            while len(vals) < c:
                if not buf:
                    buf = self.lines[self.i].split()
                    self.i += 1
                    continue
                vals.append(int(buf.pop(0)))
            out.append(vals)
        return out


_HEADER = ("llacc lldon nbone ntap lbond lbend ltors litor lpnbd int14p int13p nspw nunit lphyd ngrp nrigg "
           "nrigat nmapnl nmol nato_slv adihed slv_exist slt_exist").split()


def read_tpgprm(path, amphi=False):
    """Topology-parameter dump, readtp.f:45-150 (the non-amphi potential record
    unless amphi).  Arrays come back 1-based as stored, Fortran layouts as
    (rows, n): grppt (2, ngrp), lbnd (2, lbond), lbndg (3, lbend), int14 (2, int14p),
    int13 (2, int13p), ltor (4, ltors); potbo (lbond, 2) and potbe (lbend, 7) as
    mts_intra_n0.f:73-85 passes them (fpbond: rest length potbo(:,2), constant
    potbo(:,1); fpbend: potbe columns 2, 1, 4, 3, 5, 6, 7); litr (4, litor) and
    potit (litor, 3) (fitors: columns 2, 1, 3, mts_intra_n0.f:97-98); potto is (ltors, 2): barrier, then the signed
    multiplicity (fptors receives potto(1,2) as pota and potto(1,1) as potb,
    mts_intra_n1.f:118-119).  Charges are in electrons, masses in amu."""
    r = _ListDirected(path)
    h = dict(zip(_HEADER, r.values(len(_HEADER))))
    n = {k: int(v) for k, v in h.items() if v not in ("T", "F")}
    t = dict(header=n)
    sp = r.ints(2)
    r.ints(int(sp[0] + sp[1]))                                     # ss_point lists
    t["ss_index"] = r.ints(n["ntap"])
    r.ints(2 * n["llacc"] + 2 * n["lldon"])                        # acceptors, donors
    t["nbtype"] = r.ints(n["ntap"])
    r.ints(n["ntap"])                                              # nres(:, 1)
    r.ints(n["ntap"])                                              # nres(:, 2)
    r.ints(n["ntap"])                                              # mback
    t["grppt"] = r.ints(2 * n["ngrp"]).reshape(2, n["ngrp"])
    t["lbnd"] = r.ints(2 * n["lbond"]).reshape(2, n["lbond"])
    r.counted_lists(n["ntap"])                                     # concta
    t["lbndg"] = r.ints(3 * n["lbend"]).reshape(3, n["lbend"])
    t["ltor"] = r.ints(4 * n["ltors"]).reshape(4, n["ltors"])      # ((ltor(j,i),i=1,ltors),j=1,4)
    t["litr"] = r.ints(4 * n["litor"]).reshape(4, n["litor"])   # improper torsions, as ltor
    v = r.ints(3 * n["int14p"] + 2 * n["int13p"] + n["lpnbd"] ** 2)
    t["int14"] = v[:2 * n["int14p"]].reshape(n["int14p"], 2).T
    t["type14"] = v[2 * n["int14p"]:3 * n["int14p"]]
    o = 3 * n["int14p"]
    t["int13"] = v[o:o + 2 * n["int13p"]].reshape(n["int13p"], 2).T
    q = r.reals(1 + 2 * n["ntap"])
    t["chrge"], t["mass"] = q[1:1 + n["ntap"]], q[1 + n["ntap"]:]
    t["mapnl"] = r.ints(n["nmapnl"])
    nbun, iz = r.ints(2)
    t["iz"] = int(iz)
    L = n["lpnbd"]
    npot = 2 * n["lbond"] + 7 * n["lbend"] + (4 if amphi else 2) * n["ltors"] + 3 * n["litor"] + \
        (6 if amphi else 4) * L + 2 * n["lphyd"]
    p = r.reals(npot)
    t["potbo"] = p[:2 * n["lbond"]].reshape(2, n["lbond"]).T      # potbo(lbond, 2): constant, rest length
    o = 2 * n["lbond"]                                             # potbe(lbend, 7) (readtp.f:199-200)
    t["potbe"] = p[o:o + 7 * n["lbend"]].reshape(7, n["lbend"]).T
    if not amphi:                                                  # potto(ltors, 2): barrier, multiplicity
        o = 2 * n["lbond"] + 7 * n["lbend"]
        t["potto"] = p[o:o + 2 * n["ltors"]].reshape(2, n["ltors"]).T
    o = 2 * n["lbond"] + 7 * n["lbend"] + (4 if amphi else 2) * n["ltors"]
    t["potit"] = p[o:o + 3 * n["litor"]].reshape(3, n["litor"]).T  # potit(litor, 3)
    o = 2 * n["lbond"] + 7 * n["lbend"] + (4 if amphi else 2) * n["ltors"] + 3 * n["litor"]
    t["pnbd"] = p[o:o + 4 * L].reshape(4, L)                       # pnbd1..4 (readtp.f:140-143)
    if iz == 1:
        e = r.reals(L * (L + 1))
        t["ecc6"], t["ecc12"] = e[:L * (L + 1) // 2], e[L * (L + 1) // 2:]
    return t


def read_restart(path):
    """ASCII restart, readrs.f:216-233 (no replica-exchange header): the cell,
    step and temperature, atom positions (Cartesian, A, as the integrator
    holds them when it dumps) and velocities; the group and molecule centres
    that follow are recomputed by the path and not returned."""
    r = read_restart_full(path, positions_only=True)
    return {k: r[k] for k in ("co", "oc", "nstep", "temp", "x", "v", "ngrp", "nprot")}


def _logical(tok):
    t = tok.strip().upper().lstrip(".")
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ValueError(f"not a Fortran logical: {tok!r}")


def read_restart_full(path, positions_only=False):
    """Every record of an ASCII restart in the order dumprs.f:118-147 writes
    them (no replica-exchange or tempering header): co and oc row by row;
    nstep, temp; ntap, ngrp, nprot; positions, velocities, group centres,
    molecule centres and their velocities; the previous velocities vpx1..;
    na1, na2 and the running averages; the hoover and cpress flags, with the
    thermostat (nh; gh, vh, vh1) and barostat (zz, zz1 row by row, lzz, lzz1)
    records they switch on; igr, with igrn and krdf.  positions_only stops
    after the positions, velocities and centres."""
    r = _ListDirected(path)
    m = r.reals(18)
    co = m[:9].reshape(3, 3)            # written as ((co(i,j), j=1,3), i=1,3): row by row
    oc = m[9:].reshape(3, 3)
    nstep, temp = r.values(2)
    ntap, ngrp, nprot = (int(x) for x in r.values(3))
    v = r.reals(6 * ntap + 3 * ngrp + 6 * nprot)
    o = 6 * ntap
    out = dict(co=np.asfortranarray(co), oc=np.asfortranarray(oc), nstep=int(float(nstep)), temp=float(temp),
               x=v[:3 * ntap].reshape(3, ntap), v=v[3 * ntap:6 * ntap].reshape(3, ntap), ngrp=ngrp, nprot=nprot,
               xpg=v[o:o + 3 * ngrp].reshape(3, ngrp), xcm=v[o + 3 * ngrp:o + 3 * ngrp + 3 * nprot].reshape(3, nprot),
               vcm=v[o + 3 * ngrp + 3 * nprot:].reshape(3, nprot))
    if positions_only:
        return out
    out["v1"] = r.reals(3 * ntap).reshape(3, ntap)
    na1, na2 = (int(x) for x in r.values(2))
    avg = r.reals(na1 + na2)
    out.update(avg1=avg[:na1], avg2=avg[na1:])
    hoover, cpress = (_logical(t) for t in r.values(2))
    out.update(hoover=hoover, cpress=cpress)
    if hoover:
        nh = int(r.values(1)[0])
        h = r.reals(3 * nh)
        out.update(nh=nh, gh=h[:nh], vh=h[nh:2 * nh], vh1=h[2 * nh:])
    if cpress:
        z = r.reals(20)
        out.update(zz=z[:9].reshape(3, 3), zz1=z[9:18].reshape(3, 3), lzz=float(z[18]), lzz1=float(z[19]))
    igr = _logical(r.values(1)[0])
    out["igr"] = igr
    if igr:
        igrn = int(r.values(1)[0])
        out.update(igrn=igrn, krdf=r.ints(igrn))
    out["complete"] = all(not ln.strip() for ln in r.lines[r.i:])  # nothing left after the last record
    return out


def _record(vals, per_line=4):
    """One list-directed WRITE: values on lines of per_line (a READ takes as
    many lines as it needs, and the next record starts on a new line)."""
    toks = []
    for v in vals:
        if isinstance(v, (bool, np.bool_)):
            toks.append("T" if v else "F")
        elif isinstance(v, (int, np.integer)):
            toks.append(str(int(v)))
        else:
            toks.append(repr(float(v)))  # shortest string that reads back to the same double
    if not toks:
        return "\n"
    return "\n".join(" " + " ".join(toks[i:i + per_line]) for i in range(0, len(toks), per_line)) + "\n"


def write_restart(path, rst):
    """Write an ASCII restart with the records of dumprs.f:118-147 (serial,
    no tempering header) from a dict in read_restart_full's layout: co, oc,
    nstep, temp, x and v (3, ntap), xpg (3, ngrp), xcm and vcm (3, nprot);
    optional v1 (default v), avg1, avg2, hoover (nh, gh, vh, vh1), cpress (zz,
    zz1, lzz, lzz1), igr (krdf).  Reals are written as the shortest decimal
    that reads back to the same double, so a write and a read are exact."""
    x, v = np.asarray(rst["x"], float), np.asarray(rst["v"], float)
    xpg, xcm, vcm = (np.asarray(rst[k], float).reshape(3, -1) for k in ("xpg", "xcm", "vcm"))
    ntap, ngrp, nprot = x.shape[1], xpg.shape[1], xcm.shape[1]
    co, oc = np.asarray(rst["co"], float), np.asarray(rst["oc"], float)
    v1 = np.asarray(rst.get("v1", v), float)
    avg1, avg2 = np.asarray(rst.get("avg1", ()), float), np.asarray(rst.get("avg2", ()), float)
    hoover, cpress, igr = bool(rst.get("hoover", False)), bool(rst.get("cpress", False)), bool(rst.get("igr", False))
    parts = [
        _record(list(co.ravel(order="C")) + list(oc.ravel(order="C"))),
        _record([int(rst["nstep"]), float(rst["temp"])]),
        _record([ntap, ngrp, nprot]),
        _record(list(x[0]) + list(x[1]) + list(x[2]) + list(v[0]) + list(v[1]) + list(v[2]) + list(xpg[0])
                + list(xpg[1]) + list(xpg[2]) + list(xcm[0]) + list(xcm[1]) + list(xcm[2]) + list(vcm[0])
                + list(vcm[1]) + list(vcm[2])),
        _record(list(v1[0]) + list(v1[1]) + list(v1[2])),
        _record([len(avg1), len(avg2)]),
        _record(list(avg1) + list(avg2)),
        _record([hoover, cpress]),
    ]
    if hoover:
        gh, vh, vh1 = (np.asarray(rst[k], float) for k in ("gh", "vh", "vh1"))
        parts += [_record([len(gh)]), _record(list(gh) + list(vh) + list(vh1))]
    if cpress:
        zz, zz1 = np.asarray(rst["zz"], float), np.asarray(rst["zz1"], float)
        parts.append(_record(list(zz.ravel(order="C")) + list(zz1.ravel(order="C")) + [float(rst["lzz"]),
                                                                                       float(rst["lzz1"])]))
    parts.append(_record([igr]))
    if igr:
        krdf = np.asarray(rst["krdf"], int)
        parts += [_record([len(krdf)]), _record(list(krdf))]
    with open(path, "w") as f:
        f.write("".join(parts))


def restart_state(s, x, v, nstep, temp):
    """The restart records of an NVE state of system s: Cartesian positions x
    and velocities v (3, N) in the reference's units, the cell of s, the
    mass-weighted group centres (appbou.f) and molecule centres of mass with
    their velocities; the previous velocities are the current ones."""
    x = np.asarray(x, float)
    v = np.asarray(v, float)
    first = s.grppt[0] - 1
    last = s.grppt[1] - 1
    xpg = np.zeros((3, s.ngrp))
    for g in range(s.ngrp):
        sl = slice(first[g], last[g] + 1)
        xpg[:, g] = (x[:, sl] * s.pmass[sl]).sum(axis=1)
    nprot = int(s.atomp.max())
    m = s.mass
    mol = s.atomp - 1
    mt = np.bincount(mol, weights=m, minlength=nprot)
    xcm = np.stack([np.bincount(mol, weights=m * x[k], minlength=nprot) / mt for k in range(3)])
    vcm = np.stack([np.bincount(mol, weights=m * v[k], minlength=nprot) / mt for k in range(3)])
    return dict(co=s.co, oc=s.oc, nstep=int(nstep), temp=float(temp), x=x, v=v, xpg=xpg, xcm=xcm, vcm=vcm, v1=v,
                avg1=np.zeros(0), avg2=np.zeros(0), hoover=False, cpress=False, igr=False)


def lj_tables_amber(rmin_half, eps):
    """Packed LJ tables with the AMBER mixing rule of covnb.f:81-99:
    sigma_ij = (R_i + R_j) 2^(-1/6), eps_ij = sqrt(|eps_i eps_j|), with the
    default-real 4.184 of covnb.f:96."""
    nt = len(rmin_half)
    fact = 1.0 / (2.0 ** (1.0 / 6.0))
    npk = nt * (nt + 1) // 2
    ecc6, ecc12 = np.zeros(npk), np.zeros(npk)
    tt = np.zeros((nt, nt), dtype=np.int32)
    for i in range(nt):
        for j in range(i, nt):
            ij = (j + 1) * j // 2 + i            # 0-based packed index of (i+1, j+1)
            qener = math.sqrt(abs(eps[i] * eps[j]))
            tmpi = 4.0 * 1000.0 * qener * synth.KCAL_F32 / (synth.UNITE * synth.AVOGAD)
            tmpj = ((rmin_half[i] + rmin_half[j]) * fact) ** 6
            ecc6[ij] = tmpi * tmpj
            ecc12[ij] = tmpi * tmpj ** 2
            tt[i, j] = tt[j, i] = ij + 1
    return np.asfortranarray(tt), ecc12, ecc6


def _molecules(ntap, pairs):
    """Connected components of the bond graph, numbered by first atom (1-based)."""
    parent = np.arange(ntap)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for a, b in pairs.T - 1:
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    roots = np.array([find(a) for a in range(ntap)])
    _, first, inv = np.unique(roots, return_index=True, return_inverse=True)
    rank = np.empty(len(first), dtype=np.int32)
    rank[np.argsort(first)] = np.arange(len(first), dtype=np.int32)
    return rank[inv] + 1


def reference_system(top, rst, *, alpha, nfft, order, rcut, rtol=0.3, rnei=0.3, name=""):
    """A System of the path from a topology dump and a restart: charges to
    internal units (covchg.f:38), AMBER LJ tables, molecules from the bond
    graph, group mass fractions (redmss.f).  Per-step reciprocal corrections:
    the solute's bonds and 1-3 pairs (fudge 1, mts_furier.f:108-120); the
    solvent's bonds are its rigid constraints."""
    ntap = top["header"]["ntap"]
    ss = top["ss_index"].astype(np.int32)
    grppt = np.asfortranarray(top["grppt"].astype(np.int32))
    ngrp = grppt.shape[1]
    atomg = np.zeros(ntap, dtype=np.int32)
    for g in range(ngrp):
        atomg[grppt[0, g] - 1:grppt[1, g]] = g + 1
    if top["iz"] == 1:
        L = top["header"]["lpnbd"]
        tt = np.zeros((L, L), dtype=np.int32)
        for i in range(L):
            for j in range(i, L):
                tt[i, j] = tt[j, i] = (j + 1) * j // 2 + i + 1
        f = 1000.0 * synth.KCAL_F32 / (synth.UNITE * synth.AVOGAD)
        tt, ecc12, ecc6 = np.asfortranarray(tt), top["ecc12"] * f, top["ecc6"] * f
    else:
        tt, ecc12, ecc6 = lj_tables_amber(top["pnbd"][0], top["pnbd"][1])
    mass = top["mass"].astype(float)
    gm = np.zeros(ngrp)
    np.add.at(gm, atomg - 1, mass)
    pmass = mass / gm[atomg - 1]
    lbnd = top["lbnd"].astype(np.int32)
    atomp = _molecules(ntap, lbnd)
    solute = ss[lbnd[0] - 1] == 1
    co, oc = rst["co"], rst["oc"]
    volume = 8.0 * float(np.linalg.det(co))
    xp = oc @ rst["x"]  # reduced coordinates
    return synth.System(L=2.0 * float(co[0, 0]), xp0=np.ascontiguousarray(xp[0]),
                        yp0=np.ascontiguousarray(xp[1]), zp0=np.ascontiguousarray(xp[2]),
                        charge=synth.charge_internal(top["chrge"]), mass=mass, nbtype=top["nbtype"].astype(np.int32),
                        ntypes=int(tt.shape[0]), type_table=tt, ecc12=ecc12, ecc6=ecc6, grppt=grppt, atomg=atomg,
                        atomp=atomp, ss_index=ss, pmass=pmass, co=np.asfortranarray(co), oc=np.asfortranarray(oc),
                        volume=volume, alpha=alpha, nfft=tuple(nfft), order=order, rcut=rcut, rtol=rtol, rnei=rnei,
                        constraints=np.asfortranarray(lbnd[:, ~solute]),
                        bonds=np.asfortranarray(lbnd[:, solute]),
                        bends=np.asfortranarray(top["int13"].astype(np.int32)), name=name)


def system_from_arrays(a, name=""):
    """A System from the arrays of a stored system (e.g. tests/golden/fkbp12.npz):
    coordinates, charges (internal units), masses, LJ tables and topology, with
    params = [alpha, rcut, rtol, rnei, order, ncell, K1, K2, K3]."""
    p = a["params"]
    grppt = np.asfortranarray(a["grppt"].astype(np.int32))
    ntap = a["xp0"].shape[0]
    atomg = np.zeros(ntap, dtype=np.int32)
    for g in range(grppt.shape[1]):
        atomg[grppt[0, g] - 1:grppt[1, g]] = g + 1
    mass = a["mass"].astype(float)
    gm = np.zeros(grppt.shape[1])
    np.add.at(gm, atomg - 1, mass)
    co = np.asfortranarray(a["co"])
    return synth.System(L=2.0 * float(co[0, 0]), xp0=np.ascontiguousarray(a["xp0"]),
                        yp0=np.ascontiguousarray(a["yp0"]), zp0=np.ascontiguousarray(a["zp0"]),
                        charge=np.ascontiguousarray(a["charge"]), mass=mass, nbtype=a["nbtype"].astype(np.int32),
                        ntypes=int(a["type_table"].shape[0]), type_table=np.asfortranarray(a["type_table"].astype(
                            np.int32)), ecc12=np.ascontiguousarray(a["ecc12"]), ecc6=np.ascontiguousarray(a["ecc6"]),
                        grppt=grppt, atomg=atomg, atomp=a["atomp"].astype(np.int32),
                        ss_index=a["ss_index"].astype(np.int32), pmass=mass / gm[atomg - 1], co=co,
                        oc=np.asfortranarray(a["oc"]), volume=8.0 * float(np.linalg.det(co)), alpha=float(p[0]),
                        nfft=(int(p[6]), int(p[7]), int(p[8])), order=int(p[4]), rcut=float(p[1]), rtol=float(p[2]),
                        rnei=float(p[3]), constraints=np.asfortranarray(a["constraints"].astype(np.int32)),
                        bonds=np.asfortranarray(a["bonds"].astype(np.int32)),
                        bends=np.asfortranarray(a["bends"].astype(np.int32)), name=name)


FKBP12_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))), "tests",
                           "golden", "fkbp12.npz")


def fkbp12(path=None):
    """The reference's test1 system (tests/golden/fkbp12.npz) and its linked-cell count."""
    with np.load(path or FKBP12_PATH) as a:
        return system_from_arrays(a, name="fkbp12"), int(a["params"][5])
