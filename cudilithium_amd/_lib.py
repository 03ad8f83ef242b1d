"""ctypes loader for libcudilithium_amd.so (the C-ABI in include/cudilithium_amd.h).

The product path is the HIP library only: if the shared object is missing this
module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
# CUDILITHIUM_AMD_LIB: another build of the library (for example one compiled
# with different tuning flags by build.build(out=..., extra_flags=...))
LIB_PATH = os.environ.get("CUDILITHIUM_AMD_LIB") or os.path.join(HERE, "libcudilithium_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cudilithium_amd.h")

c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_int = ctypes.c_int
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_szp = ctypes.POINTER(ctypes.c_size_t)


class DilError(RuntimeError):
    pass


_lib = None


def declared_symbols():
    """Function names declared in include/cudilithium_amd.h (and api headers)."""
    names = []
    inc = os.path.dirname(HEADER)
    for root, _, files in os.walk(inc):
        for f in files:
            if f.endswith(".h"):
                txt = open(os.path.join(root, f)).read()
                txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
                names += re.findall(r"\b(dil\w*)\s*\(", txt)
    return sorted(set(names))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DilError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                       "python cudilithium_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "dil_public_key_bytes": (c_size_t, [c_int]),
        "dil_secret_key_bytes": (c_size_t, [c_int]),
        "dil_signature_bytes": (c_size_t, [c_int]),
        "dil_ntt_forward": (c_int, [c_void_p, c_size_t, c_size_t, c_void_p]),
        "dil_ntt_inverse": (c_int, [c_void_p, c_size_t, c_size_t, c_void_p]),
        "dil_ntt_pointwise_pipeline": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_void_p]),
        "dil_keccak_f1600_batch": (c_int, [c_void_p, c_size_t, c_void_p]),
        "dil_device_count": (c_int, [ctypes.POINTER(c_int)]),
        "dil_set_device": (c_int, [c_int]),
        "dil_device_name": (c_int, [ctypes.c_char_p, c_size_t]),
        "dil_device_cu_count": (c_int, []),
        "dil_malloc_device": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
        "dil_malloc_pitch": (c_int, [ctypes.POINTER(c_void_p), _szp, c_size_t, c_size_t]),
        "dil_free_device": (c_int, [c_void_p]),
        "dil_malloc_host": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
        "dil_free_host": (c_int, [c_void_p]),
        "dil_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
        "dil_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
        "dil_memcpy_d2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
        "dil_memset_device": (c_int, [c_void_p, c_int, c_size_t, c_void_p]),
        "dil_stream_create": (c_int, [ctypes.POINTER(c_void_p)]),
        "dil_stream_destroy": (c_int, [c_void_p]),
        "dil_stream_synchronize": (c_int, [c_void_p]),
        "dil_device_synchronize": (c_int, []),
        "dil_event_create": (c_int, [ctypes.POINTER(c_void_p)]),
        "dil_event_destroy": (c_int, [c_void_p]),
        "dil_event_record": (c_int, [c_void_p, c_void_p]),
        "dil_event_synchronize": (c_int, [c_void_p]),
        "dil_event_elapsed_ms": (c_int, [ctypes.POINTER(ctypes.c_float), c_void_p, c_void_p]),
        "dil_error_string": (ctypes.c_char_p, [c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int, what: str = "") -> None:
    if code != 0:
        msg = load().dil_error_string(code).decode()
        raise DilError(f"{what or 'cudilithium_amd'} failed: {msg} ({code})")


def register(name: str, restype, argtypes) -> None:
    fn = getattr(load(), name)
    fn.restype = restype
    fn.argtypes = argtypes
