"""Builds the native library of nsphere_amd in-tree for gfx950.

    python nsphere_amd/build.py          # -> nsphere_amd/libnsphere_gpu.so

hipcc compiles the HIP translation units (cross-compiles without a GPU);
the host-only IC sampler is compiled with g++ and OpenMP.  All HIP code is
built with -ffp-contract=off (the strict-IEEE integrator depends on it) and
with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libnsphere_gpu.so")
ARCH = "gfx950"

HIP_SOURCES = ["ns_api.hip", "ns_sort.hip", "ns_integrate.hip", "ns_scan.hip", "ns_mlp.hip", "ns_io.hip",
               "ns_density.hip", "ns_rank.hip", "ns_overlap.hip", "ns_noalloc.hip", "ns_qsort.hip"]
HOST_SOURCES = ["ns_ic.cpp", "ns_dropin.cpp"]
DRIVER_SRC = os.path.join(ROOT, "tests", "c", "dropin_driver.c")
DRIVER_BIN = os.path.join(ROOT, "tests", "c", "dropin_driver")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm 7.x expected under /opt/rocm)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")
    # keep real warnings visible, drop the host pass's note about the GPU feature flag
    for line in proc.stdout.splitlines():
        if "packed-fp32-ops" in line or "not a recognized feature" in line:
            continue
        if line.strip():
            sys.stderr.write(line + "\n")


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    inc = ["-I", INCLUDE, "-I", CSRC]
    with tempfile.TemporaryDirectory(prefix="nsphere_build_") as tmp:
        objs = []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            _run([hipcc, *HIPCC_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        for src in HOST_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unknown-pragmas", *inc, "-c",
                  os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        tmp_lib = LIB + ".tmp"
        _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib, "-lgomp"])
        os.replace(tmp_lib, LIB)
    if os.path.exists(DRIVER_SRC):
        # a plain C program using the reference's signatures (nsphere_dropin.h)
        _run(["gcc", "-O2", "-std=c11", "-Wall", "-I", INCLUDE, DRIVER_SRC, "-o", DRIVER_BIN,
              "-L", PKG, "-lnsphere_gpu", "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/../../nsphere_amd"])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
