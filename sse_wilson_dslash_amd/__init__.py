"""MI355X-native Wilson Dslash (drop-in for the sse_wilson_dslash C API).

The compute path is hand-written HIP for gfx950 behind a C ABI
(include/dslash_mi355x.h, include/sse_dslash.h, include/sse_dslash_3d.h);
this package binds it with ctypes and uses torch only for device memory and
streams.
"""
from ._lib import F32, F64, F16, BF16, F64S, DslashError  # noqa: F401

__all__ = ["F32", "F64", "F16", "BF16", "DslashError"]
