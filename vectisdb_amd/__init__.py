"""vectisdb_amd -- MI355X-native execution engine for VectisDB's DuckDB GPU path.

The engine itself is C++/HIP in libvgpu.so (vectisdb_amd/csrc, built by vectisdb_amd.build); this
package only holds the ctypes binding (capi) and the Python mirror of the engine's plan API (engine).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
