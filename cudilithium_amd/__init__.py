"""MI355X-native batched CRYSTALS-Dilithium (round-3 v3.1, modes 2/3/5).

The compute path is libcudilithium_amd.so (hand-written HIP kernels for
gfx950 behind the C-ABI in include/cudilithium_amd.h).  This package is the
host-side mirror of the reference's batched interface for Python callers.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
