"""Build the native parts of letkf_amd in-tree.

* ``letkf_amd/liblekf_hip.so``  -- HIP kernels + C ABI (hipcc, gfx950)
* ``letkf_amd/fortran/liblekf_fortran.so`` -- the Fortran host modules
  (``common_letkf`` drop-in and the ``das_letkf`` shims, amdflang) linked
  against liblekf_hip.so through ISO_C_BINDING.

Run ``python -m letkf_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "liblekf_hip.so")
FLIB = os.path.join(FSRC, "liblekf_fortran.so")

HIP_SOURCES = ["lekf_solve.hip", "lekf_cull.hip", "lekf_obsope.hip", "lekf_regional.hip",
               "lekf_relax.hip", "lekf_l96.hip", "lekf_radar.hip", "lekf_capi.hip"]
F_SOURCES = ["common_letkf.f90", "letkf_tools_gpu.f90", "letkf_speedy_gpu.f90"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flang() -> str | None:
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... ({r.returncode})")
    return r


def build_hip(verbose: bool = False) -> str:
    """Compile every HIP source to an object (in parallel) and link the library."""
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    cmds = [[_hipcc(), *HIPCC_FLAGS, "-I", INC, "-I", CSRC, "-c", s, "-o", o]
            for s, o in zip(srcs, objs)]
    jobs = max(1, min(len(cmds), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([_hipcc(), *HIPCC_FLAGS, "-shared", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    fc = _flang()
    if fc is None:
        return None
    srcs = [os.path.join(FSRC, s) for s in F_SOURCES if os.path.exists(os.path.join(FSRC, s))]
    if not srcs:
        return None
    mod = os.path.join(FSRC, "_mod")
    os.makedirs(mod, exist_ok=True)
    tmp = FLIB + ".tmp"
    cmd = [fc, "-O2", "-fPIC", "-shared", "-module-dir", mod, *srcs, "-L", HERE, "-llekf_hip",
           "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    _run(cmd, verbose)
    os.replace(tmp, FLIB)
    # the driver programs of the Fortran host side (used by the tests):
    # the Lorenz-96 loop, the SPEEDY analysis sequence and letkf_core from
    # OpenMP threads
    for name in ("l96_driver", "speedy_driver", "omp_core_driver"):
        exe = os.path.join(FSRC, name)
        omp = ["-fopenmp"] if name == "omp_core_driver" else []
        cmd = [fc, "-O2", *omp, "-I", mod, os.path.join(FSRC, name + ".f90"), "-L", FSRC,
               "-llekf_fortran", "-L", HERE, "-llekf_hip", "-Wl,-rpath,$ORIGIN",
               "-Wl,-rpath,$ORIGIN/..", "-o", exe + ".tmp"]
        _run(cmd, verbose)
        os.replace(exe + ".tmp", exe)
    return FLIB


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
