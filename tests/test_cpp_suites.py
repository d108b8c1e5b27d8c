"""The C++ drop-in suites: tests/cpp/*.cpp compiled against include/*.hpp and linked to libvgpu.so the way
the DuckDB extension links the engine (built by tests/cpp/build.py, which __graft_entry__.build() runs)."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "_bin"


def run_suite(name, timeout):
    exe = BIN / name
    if not exe.exists():
        import importlib.util
        spec = importlib.util.spec_from_file_location("cpp_build", Path(__file__).resolve().parent / "cpp" / "build.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=timeout)
    tail = (res.stdout[-3000:] + "\n" + res.stderr[-3000:]).strip()
    assert res.returncode == 0, tail
    assert "expectations met" in res.stdout and "not ok" not in res.stdout, tail


def test_host_components_suite():
    """KernelCache + SHA-256, GpuLogger and the routing functions through the C++ headers (no GPU)."""
    run_suite("test_host_components", 120)


@pytest.mark.gpu
def test_device_components_suite():
    """FilterGreaterThanInt32, CodeGenerator, RmmPool, GpuMemoryPool and GpuEngine through the C++ headers."""
    run_suite("test_device_components", 120)


@pytest.mark.gpu
def test_stream_pipeline_suite():
    """GpuStreamPipeline: overlap, FIFO order, chained levels, NULLs, slot release, RUNNING_SUM, wide levels."""
    run_suite("test_stream_pipeline", 120)
