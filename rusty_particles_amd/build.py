"""Build the native libraries in-tree.

librp_hip.so   HIP engine for gfx950 + C ABI (include/rp_abi.h)
librp_scene.so host-only scene generators used by bench.py and tests

Run as ``python -m rusty_particles_amd.build``.  The built files sit next to
this file so they travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIP_LIB = os.path.join(PKG, "librp_hip.so")
SCENE_LIB = os.path.join(PKG, "librp_scene.so")

# -ffp-contract=off: no fused multiply-add, so the pair predicate and every
# contact force equal a plain IEEE evaluation of the reference expression.
# -packed-fp32-ops off: no packed FP32 instructions are emitted.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> str:
    tmp = HIP_LIB + ".tmp"
    srcs = [os.path.join(CSRC, "rp_engine.hip"), os.path.join(CSRC, "rp_surrogate.hip"),
            os.path.join(CSRC, "rp_stl.cpp")]
    _run([_hipcc(), *HIP_FLAGS, *srcs, "-o", tmp])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_scene() -> str:
    tmp = SCENE_LIB + ".tmp"
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
          os.path.join(CSRC, "rp_scene.cpp"), "-o", tmp])
    os.replace(tmp, SCENE_LIB)
    return SCENE_LIB


def build_all() -> None:
    build_hip()
    build_scene()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "scene": build_scene}[w]()
