"""MI355X-native particle-filter step of LibBi (lawmurray/LibBi v1.4.5).

The product is the C ABI in include/bihip.h, implemented by hand-written
gfx950 HIP kernels in libbi_amd/csrc and built into libbi_amd/lib/libbihip.so;
the C++ shims in include/bi/amd/ keep the reference's template call sites.
This package holds the build step (libbi_amd.build), a ctypes binding of the
C ABI (libbi_amd.capi) and synthetic-data helpers (libbi_amd.data).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
