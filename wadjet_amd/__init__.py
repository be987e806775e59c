"""wadjet_amd — MI355X-native hash aggregate / hash join for wadjet's
vectorized engine path (HIP kernels behind a C-ABI, see include/wadjet_gpu.h).
"""
from . import _lib  # noqa: F401
from ._lib import lib, WgError, NullKeyError, UnsupportedError  # noqa: F401

__all__ = ["lib", "WgError", "NullKeyError", "UnsupportedError"]
