"""MI355X (gfx950) path-tracing backend for Aurora.

The product is libaph.so (hand-written HIP kernels behind the C-ABI of
include/aph.h).  This package holds its ctypes binding (capi), a thin host
wrapper (session.Renderer), the seeded synthetic workloads (scenes) and the
build recipe (build).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
