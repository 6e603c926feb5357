"""MI355X (gfx950) N-body engine: the force-and-step path of gtd::system.

The engine is libgregnbody.so (HIP kernels behind the C ABI of
include/gregnbody.h); `capi` binds it for Python callers and tests.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
