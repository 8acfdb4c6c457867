"""Builds libswf.so (HIP, gfx950) and the Fortran driver in-tree.

    python -m sweeping_fd_weno_amd.build          # library + driver
Every HIP translation unit is compiled for gfx950 with packed-FP32 instructions
disabled.  The FP64-exact and mode-independent kernels (swf_k_exact.hip) and the
host API are compiled without floating-point contraction, the FP64/FP32-bulk
kernels (swf_k_fast.hip) with contraction.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libswf.so")
OBJ = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = os.environ.get("SWF_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-fast-math", "-I", os.path.join(ROOT, "include"), "-Wno-unused-result",
]
UNITS = [
    ("swf_k_exact.hip", ["-ffp-contract=off"]),
    ("swf_k_fast.hip", ["-ffp-contract=fast"]),
    ("swf_api.hip", ["-ffp-contract=off"]),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    procs = []
    for src, extra in UNITS:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), cmd))
        objs.append(obj)
    for p, cmd in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + out + err)
            raise RuntimeError("hipcc failed on " + cmd[-3])
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> list:
    """The Fortran drivers (reference parameter surfaces) linked against libswf.so:
    swf_driver (ex4 Sedov), swf_vortex (ex3 vortex accuracy), swf_sedov1d (ex2 1D Sedov) and swf_jet
    (ex5 Mach 2000 jet)."""
    drv_dir = os.path.join(PKG, "fortran")
    mod_src = os.path.join(drv_dir, "swf_mod.f90")
    if not os.path.exists(FLANG) or not os.path.exists(mod_src):
        return []
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    built = []
    for prog in ("swf_driver", "swf_vortex", "swf_sedov1d", "swf_jet"):
        src = os.path.join(drv_dir, prog + ".f90")
        if not os.path.exists(src):
            continue
        out = os.path.join(drv_dir, prog)
        _run([FLANG, "-O2", "-fdefault-real-8", "-module-dir", mod, mod_src, src, "-o", out,
              "-L" + PKG, "-lswf", "-Wl,-rpath,$ORIGIN/.."])
        built.append(out)
    return built


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_driver()


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
