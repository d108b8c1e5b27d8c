"""Builds the C++ drop-in suites (tests/cpp/_bin/) against include/*.hpp and vectisdb_amd/lib/libvgpu.so,
the way the DuckDB extension compiles and links against the engine. Run by __graft_entry__.build()."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
BIN = HERE / "_bin"
RPATH = "-Wl,-rpath,$ORIGIN/../../../vectisdb_amd/lib"


def build() -> list:
    BIN.mkdir(exist_ok=True)
    common = ["-std=c++17", "-O1", f"-I{ROOT / 'include'}", f"-L{ROOT / 'vectisdb_amd' / 'lib'}", "-lvgpu", RPATH]
    host = ["g++", str(HERE / "test_host_components.cpp"), "-o", str(BIN / "test_host_components"), *common,
            "-pthread"]
    device = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", str(HERE / "test_device_components.cpp"), "-o",
              str(BIN / "test_device_components"), *common, "-pthread"]
    pipeline = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", str(HERE / "test_stream_pipeline.cpp"), "-o",
                str(BIN / "test_stream_pipeline"), *common, "-pthread"]
    for cmd in (host, device, pipeline):
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"build failed:\n{' '.join(cmd)}\n{res.stderr}")
    return [BIN / "test_host_components", BIN / "test_device_components", BIN / "test_stream_pipeline"]


if __name__ == "__main__":
    for p in build():
        print(p)
