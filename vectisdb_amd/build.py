"""Build recipe for libvgpu.so (the HIP kernels, the C++ engine and the extern "C" boundary).

Every source under vectisdb_amd/csrc is compiled with hipcc for gfx950 and linked in-tree into
vectisdb_amd/lib/libvgpu.so, so the shared object travels with the source tree. Run
`python -m vectisdb_amd.build` (or __graft_entry__.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = ROOT / "_build" / "obj"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libvgpu.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VGPU_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-command-line-argument",
    "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
    f"-I{ROOT / 'include'}",
]
# Device code: IEEE without contraction (bit parity with the CPU for +,-,*,/), and no packed FP32 ops.
DEVICE = [
    f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _sources():
    hip = sorted(CSRC.rglob("*.hip"))
    cpp = sorted(CSRC.rglob("*.cpp"))
    return hip, cpp


def _compile(src: Path) -> Path:
    rel = src.relative_to(CSRC)
    obj = OBJ / (str(rel).replace("/", "__") + ".o")
    obj.parent.mkdir(parents=True, exist_ok=True)
    if src.suffix == ".hip":
        cmd = [HIPCC, "-c", *COMMON, *DEVICE, "-x", "hip", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, "-c", *COMMON, *DEVICE, str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        noise = ("unknown target feature", "packed-fp32")
        err = "\n".join(l for l in res.stderr.splitlines() if not any(n in l for n in noise))
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{err}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    hip, cpp = _sources()
    srcs = hip + cpp
    jobs = jobs or min(8, os.cpu_count() or 4)
    if OBJ.exists():
        shutil.rmtree(OBJ)
    OBJ.mkdir(parents=True)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    LIBDIR.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs),
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    # the benchmark's GpuStreamPipeline program: a C++ caller of the drop-in headers, linked like the extension
    exe = LIBDIR / "bench_stream_pipeline"
    cmd = ["g++", str(PKG / "bench_programs" / "stream_pipeline.cpp"), "-o", str(exe), "-std=c++17", "-O2",
           f"-I{ROOT / 'include'}", f"-L{LIBDIR}", "-lvgpu", "-Wl,-rpath,$ORIGIN", "-pthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build failed:\n{' '.join(cmd)}\n{res.stderr}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, jobs=int(sys.argv[1]) if len(sys.argv) > 1 else None)
