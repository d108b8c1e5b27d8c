"""Builds oracle/_ref/libvref.so from the reference's own device-independent sources. TEST INFRASTRUCTURE.

Compiles, unmodified and in place, cuda_engine/src/gpu_cost_model.cpp and cuda_engine/src/kernel_cache.cpp
of the reference tree (they include only the reference's gpu_engine.hpp / kernel_cache.hpp and standard
headers), together with oracle/ref_shim.cpp, which exposes them to ctypes. Nothing is copied into this
repository; the output lands in oracle/_ref/ (git-ignored, but it travels with the tree). Only the tests
load it, as a checker. The rest of the reference's engine needs nvcc, NVRTC and Thrust and is not built.

Run: python -m oracle.build_ref  (needs /root/reference; __graft_entry__.build() calls it when present)
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path("/root/reference/cuda_engine")
OUT = HERE / "_ref" / "libvref.so"
SOURCES = [REF / "src" / "gpu_cost_model.cpp", REF / "src" / "kernel_cache.cpp", HERE / "ref_shim.cpp"]


def available() -> bool:
    return all(p.exists() for p in SOURCES)


def build() -> Path:
    if not available():
        raise FileNotFoundError("reference sources not present; the prebuilt oracle/_ref/libvref.so is used")
    OUT.parent.mkdir(exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-I{REF / 'include'}", "-o", str(OUT),
           *map(str, SOURCES)]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
