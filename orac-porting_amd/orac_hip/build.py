"""Build of the gfx950 shared library liborac_hip.so (in-tree).

``python -m orac_hip.build`` (with ``orac-porting_amd`` on sys.path) or
``build()`` compiles every HIP source for gfx950 with hipcc into
``orac-porting_amd/lib/liborac_hip.so``, and the Fortran drop-in shim
(``fortran/orac_hip_shim.f90``) into ``lib/liborac_hip_fortran.so`` when
flang is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                      # orac-porting_amd/
CSRC = os.path.join(ROOT, "csrc")
LIBDIR = os.path.join(ROOT, "lib")
LIB = os.path.join(LIBDIR, "liborac_hip.so")
FLIB = os.path.join(LIBDIR, "liborac_hip_fortran.so")
SOURCES = ["pme.hip", "pairs.hip", "force_pairs.hip", "bonded.hip", "md.hip", "host.cpp", "api.cpp", "engine.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ORAC_HIP_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed-FP32 instructions stay off: hipcc would otherwise emit them from float2 code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ... (exit {r.returncode})")
    return r


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    # every translation unit launches only its own kernels, so the objects
    # compile independently (in parallel) and link without relocatable device code
    cmds = [[HIPCC, *HIP_FLAGS, "-c", "-x", "hip", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as pool:
        list(pool.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    flang = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    shim = os.path.join(ROOT, "fortran", "orac_hip_shim.f90")
    if os.path.exists(shim) and os.path.exists(flang):
        mod = os.path.join(LIBDIR, "fmod")
        os.makedirs(mod, exist_ok=True)
        _run([flang, "-O2", "-cpp", "-fPIC", "-shared", "-module-dir", mod, shim, "-o", FLIB + ".tmp", f"-L{LIBDIR}",
              "-lorac_hip", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(FLIB + ".tmp", FLIB)
        # the Fortran callers used by the integration tests (tests/fortran)
        for name in ("drive_shim.f90", "drive_intra.f90"):
            drv = os.path.join(os.path.dirname(ROOT), "tests", "fortran", name)
            if os.path.exists(drv):
                exe = drv[:-4]
                _run([flang, "-O2", drv, "-o", exe + ".tmp", f"-L{LIBDIR}", "-lorac_hip_fortran", "-lorac_hip",
                      "-Wl,-rpath,$ORIGIN/../../orac-porting_amd/lib"], verbose)
                os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
