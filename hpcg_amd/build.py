"""Builds hpcg_amd/libhpcgx.so (HIP, gfx950) and the C++ drop-in objects.

    python -m hpcg_amd.build            # library only
    python -m hpcg_amd.build --dropin   # also the C++ hook library

Every HIP file is compiled with -ffp-contract=off (products and sums rounded
separately, as in the reference's x86-64 build, which the parity tests
require) and without packed fp32 instructions.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
DROPIN = os.path.join(PKG, "dropin")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhpcgx.so")
DROPIN_LIB = os.path.join(PKG, "libhpcg_dropin.so")

ARCH = os.environ.get("HPCGX_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}",
]

SOURCES = ["setup.hip", "ops.hip", "cg.hip"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, "-c", *HIP_FLAGS, "-o", obj, os.path.join(CSRC, src)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_dropin(verbose=False):
    """The reference hook signatures (CG, ComputeSPMV, ...) over the C ABI,
    as a shared library a reference build links instead of its hook files."""
    srcs = sorted(f for f in os.listdir(DROPIN) if f.endswith(".cpp"))
    if not srcs:
        return None
    tmp = DROPIN_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++11", "-fPIC", "-shared", "-Wall", "-DHPCG_NO_MPI",
          f"-I{DROPIN}", f"-I{INCLUDE}", "-o", tmp, *[os.path.join(DROPIN, s) for s in srcs],
          f"-L{PKG}", "-lhpcgx", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DROPIN_LIB)
    if verbose:
        print("built", DROPIN_LIB)
    return DROPIN_LIB


if __name__ == "__main__":
    build_library(verbose=True)
    if "--dropin" in sys.argv:
        build_dropin(verbose=True)
