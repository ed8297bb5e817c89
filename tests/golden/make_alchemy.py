"""Reference results for the alchemical routines on a seeded synthetic system.

The system is orac_hip.synth.small_solvated() (40-atom solute in 380 waters)
with its first 12 solute atoms removed alchemically (synth.alchemy_for).  For
each lambda state below the script runs the reference Fortran compiled in
place (oracle/_ref): force_alchemy and force_alchemy_shift (serial build) on
the single outer shell and on a switched inner + outer shell pair, ferrf_alchemy on the solute's
bond and 1-3 pairs (iz = 1 and 0), cself_alchemy, and fft_pme_alchemy
(fftpack branch), and writes tests/golden/alchemy.npz: energies, dwrk and
stress as values, forces of the direct routines as SHA-256 digests of their
bytes, PME forces as values.

    python tests/golden/make_alchemy.py
"""
import dataclasses
import hashlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "orac-porting_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from orac_hip import synth  # noqa: E402

E6 = ("uconf_slt", "uconf_slv", "uconf_ss", "ucoul_slt", "ucoul_slv", "ucoul_ss")
# (name, lrm, lrm0, lrmq, lrmq0, intramol, ljintra)
STATES = [
    ("full", 0.0, 0.0, 0.0, 0.0, False, False),
    ("mid", 0.35, 0.4, 0.6, 0.55, False, False),
    ("intra", 0.5, 0.45, 0.3, 0.35, True, False),
    ("ljintra", 0.7, 0.75, 0.8, 0.75, True, True),
    ("off", 1.0, 0.9, 1.0, 0.9, False, False),
]
SHELLS = [("outer", dict(rinn=0.0, rtolinn=0.0, rtolout=0.3, rneigh=0.0)),
          ("switched", dict(rinn=4.0, rout=7.0, rtolinn=0.5, rtolout=0.5, rneigh=0.6))]
N_ALCH = 12


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


def system():
    s = synth.small_solvated()
    return s, np.arange(N_ALCH)


def outer_list(s, fa):
    rc = s.rcut + s.rtol + s.rnei
    u = O.force(s, O.all_groups_list(s.ngrp), rinn=rc, rout=rc, rtolinn=0.0, rtolout=0.0, rneigh=0.0, iz=0, fa=fa)
    return u["nnlpp1"]


def state(name):
    for st in STATES:
        if st[0] == name:
            _, lrm, lrm0, lrmq, lrmq0, intra, ljin = st
            return synth.lambda_state(lrm, lrm0=lrm0, lrmq=lrmq, lrmq0=lrmq0), intra, ljin
    raise KeyError(name)


def main():
    assert O.ref_available(), "build the reference first: make -C oracle ref"
    s, atoms = system()
    fa = O.ForceArgs(s)
    nn = outer_list(s, fa)
    out = {"inputs_sha": digest(np.concatenate([s.xp0, s.yp0, s.zp0, s.charge])), "nnlpp_sha": digest(nn)}
    pairs = np.concatenate([s.bonds, s.bends], axis=1)
    for name, *_ in STATES:
        lam, intra, ljin = state(name)
        al = synth.alchemy_for(s, atoms, intramol=intra, ljintra=ljin)
        for form, alf in (("force", al), ("shift", dataclasses.replace(al, shifted=True))):
            for sh, kw in SHELLS:
                kw = dict(kw)
                kw.setdefault("rout", s.rcut)
                r = O.ref_force_alchemy(s, alf, lam, nn, fa=fa, **kw)
                out[f"{form}_{name}_{sh}_e"] = np.array([r[k] for k in E6] + [r["dwrk"]])
                out[f"{form}_{name}_{sh}_stress"] = r["stress"]
                out[f"{form}_{name}_{sh}_f_sha"] = digest(r["f"])
        q = synth.alchemy_charges(s.charge, al.ss_alchemy, lam[4], lam[5])
        q0 = synth.alchemy_charges(s.charge, al.ss_alchemy, lam[6], lam[7])
        for iz in (1, 0):
            r = O.ref_ferrf_alchemy(s, q, q0, pairs, iz, 0.8)
            out[f"ferrf_{name}_{iz}_e"] = np.array([r["slt"], r["slv"], r["dwrk"]])
            out[f"ferrf_{name}_{iz}_f_sha"] = digest(r["f"])
        out[f"cself_{name}"] = np.array(O.ref_cself_alchemy(s, q, q0))
        p = O.ref_fft_pme_alchemy(s, q, q0)
        out[f"pme_{name}_e"] = np.array([p["eer"], p["dwrk"]])
        out[f"pme_{name}_stress"] = p["stress"]
        out[f"pme_{name}_f"] = p["f"]
    np.savez_compressed(os.path.join(HERE, "alchemy.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
