"""The C ABI library loads on a host without a GPU and exports every function include/vgpu.h declares.
No compute calls are made here."""
import ctypes
import re
from pathlib import Path

from vectisdb_amd import capi

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "vgpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(vgpu_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_functions():
    names = declared_functions()
    assert "vgpu_group_by" in names and "vgpu_filter_project" in names and len(names) >= 10


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(capi.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(set(declared_functions()) - set(capi.exported_symbols())) == []


def test_abi_version_and_error_string_without_gpu():
    lib = capi.load()
    assert lib.vgpu_abi_version() >= 1
    assert lib.vgpu_last_error() is not None
