"""Builds the native libraries of coupe_amd in place.

* ``coupe_amd/libcoupe_hip.so`` -- the product: HIP kernels for gfx950 and
  the C ABI of include/coupe.h and include/coupe_hip.h (hipcc).
* ``coupe_amd/libcoupe_workloads.so`` -- synthetic input generators used by
  benchmarks and tests (gcc).
* ``coupe_amd/libcoupe_probe.so`` -- STREAM-style bandwidth probe reported by
  the benchmark (hipcc).

Run ``python -m coupe_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

LIB = os.path.join(HERE, "libcoupe_hip.so")
WORKLOADS = os.path.join(HERE, "libcoupe_workloads.so")
PROBE = os.path.join(HERE, "libcoupe_probe.so")

HIP_SOURCES = ["rcb_engine.hip", "rib.hip", "imbalance.hip", "mesh.hip", "grid.hip", "mesh_parts.hip", "coupe_api.cpp",
               "serve.cpp"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # f32 split arithmetic must be IEEE: no contraction, denormals kept.
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    # packed FP32 instructions are kept out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(verbose: bool = True) -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    # roctx trace ranges (rocprofiler-sdk)
    roctx = ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs, *roctx]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_probe() -> str:
    """The bandwidth probe bench.py reports beside the roofline (measurement only)."""
    tmp = PROBE + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "probe.hip")])
    os.replace(tmp, PROBE)
    return PROBE


def build_workloads() -> str:
    src = os.path.join(HERE, "workloads.c")
    tmp = WORKLOADS + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c11", "-Wall", "-o", tmp, src, "-lm"])
    os.replace(tmp, WORKLOADS)
    return WORKLOADS


def build_all() -> None:
    build_hip()
    build_probe()
    build_workloads()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "hip" in what:
        build_hip()
    if "all" in what or "probe" in what:
        build_probe()
    if "all" in what or "workloads" in what:
        build_workloads()
