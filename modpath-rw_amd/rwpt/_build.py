"""In-tree build of librwpt.so (hipcc, gfx950) and of the CPU oracle."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(PKG))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(os.path.dirname(PKG), "librwpt.so")

SOURCES = ["rwpt_api.hip", "rwpt_cells.hip", "rwpt_walk.hip", "rwpt_walk32.hip",
           "rwpt_kde.hip", "rwpt_gpkde.hip", "rwpt_export.hip", "rwpt_adjoint.hip",
           "rwpt_usg.hip"]

# The FP32-local walk has no bit-exactness contract: its translation unit
# contracts a * b + c into fused multiply-add; every other one keeps the
# reference's separate roundings.
CONTRACT = {"rwpt_walk32.hip": "-ffp-contract=fast"}

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_library(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, procs = [], []
    # the translation units compile in parallel
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        # FP64 parity: no contraction of a*b+c into fma (except CONTRACT)
        cmd = [hipcc, *HIPCC_FLAGS, CONTRACT.get(src, "-ffp-contract=off"), "-I", INCLUDE,
               "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, pr in procs:
        if pr.wait() != 0:
            raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


FLANG = "/opt/rocm/lib/llvm/bin/amdflang"
FORTRAN_DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_engine")


def build_fortran_driver():
    """Compiles the Fortran test driver of the engine (include/rwpt.f90 +
    tests/fortran/drive_engine.f90, linked against librwpt.so) where
    amdflang is present; the library is found next to the driver's tree."""
    if not os.path.exists(FLANG):
        return False
    import tempfile
    with tempfile.TemporaryDirectory() as mod:
        tmp = FORTRAN_DRIVER + ".tmp"
        _run([FLANG, "-O1", "-J", mod, "-o", tmp, os.path.join(INCLUDE, "rwpt.f90"),
              os.path.join(ROOT, "tests", "fortran", "drive_engine.f90"),
              "-L", os.path.dirname(LIB), "-lrwpt",
              "-Wl,-rpath,$ORIGIN/../../modpath-rw_amd"])
        os.replace(tmp, FORTRAN_DRIVER)
    return True


REFERENCE_SRC = "/root/reference/src"


def build_reference():
    """Compiles the reference's own Fortran engine into oracle/_ref (the
    checker of the oracle) where the reference sources are present."""
    if not os.path.isdir(REFERENCE_SRC):
        return False
    if not os.path.exists("/opt/rocm/lib/llvm/bin/amdflang"):
        return False
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={REFERENCE_SRC}"])
    return True


if __name__ == "__main__":
    build_library()
    if "--oracle" in sys.argv:
        build_oracle()
        build_reference()
