"""Host-side Python helpers of the MI355X clipped-Voronoi engine.

The engine itself is libvoronoi_hip.so (C ABI in include/vd_api.h); this
package holds its ctypes binding (`native`), the reference's file formats
(`fileio`) and seeded synthetic inputs (`meshgen`).
"""
from . import fileio, meshgen  # noqa: F401

__all__ = ["fileio", "meshgen", "native"]
