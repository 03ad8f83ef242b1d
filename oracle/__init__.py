"""ctypes wrapper for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline,
never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# compiler flags of liboracle.so (oracle/Makefile): a portable x86-64-v3
# build, since the library is compiled on one host and loaded on another
BUILD_FLAGS = "gcc -O3 -march=x86-64-v3 -fopenmp"

_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "mode", "k", "l", "eta", "tau", "beta", "gamma1", "gamma2", "omega",
        "pk_bytes", "sk_bytes", "sig_bytes", "polyeta_bytes", "polyz_bytes", "polyw1_bytes")]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_params.argtypes = [ctypes.c_int, ctypes.POINTER(Params)]
        L.oracle_shake128.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t]
        L.oracle_shake256.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t]
        L.oracle_keccakf1600.argtypes = [_u64p]
        L.oracle_montgomery_reduce.argtypes = [ctypes.c_int64]
        L.oracle_montgomery_reduce.restype = ctypes.c_int32
        for f in ("oracle_reduce32", "oracle_caddq"):
            getattr(L, f).argtypes = [ctypes.c_int32]
            getattr(L, f).restype = ctypes.c_int32
        L.oracle_power2round.argtypes = [_i32p, ctypes.c_int32]
        L.oracle_power2round.restype = ctypes.c_int32
        L.oracle_decompose.argtypes = [ctypes.c_int, _i32p, ctypes.c_int32]
        L.oracle_decompose.restype = ctypes.c_int32
        L.oracle_make_hint.argtypes = [ctypes.c_int, ctypes.c_int32, ctypes.c_int32]
        L.oracle_make_hint.restype = ctypes.c_int32
        L.oracle_use_hint.argtypes = [ctypes.c_int, ctypes.c_int32, ctypes.c_int32]
        L.oracle_use_hint.restype = ctypes.c_int32
        L.oracle_zetas.restype = _i32p
        L.oracle_ntt.argtypes = [_i32p]
        L.oracle_invntt_tomont.argtypes = [_i32p]
        L.oracle_ntt_pipeline.argtypes = [_i32p, _i32p, _i32p, ctypes.c_size_t]
        L.oracle_ntt_batch.argtypes = [_i32p, ctypes.c_size_t, ctypes.c_int]
        L.oracle_keypair_from_seed.argtypes = [ctypes.c_int, _u8p, _u8p, _u8p]
        L.oracle_keypair_counter.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint64]
        L.oracle_sign.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_size_t, _u8p]
        L.oracle_sign.restype = ctypes.c_int
        L.oracle_verify.argtypes = [ctypes.c_int, _u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, _u8p]
        L.oracle_verify.restype = ctypes.c_int
        L.oracle_keypair_batch.argtypes = [ctypes.c_int, _u8p, _u8p, _u8p, ctypes.c_size_t, ctypes.c_int]
        L.oracle_keypair_counter_batch.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint64,
                                                   ctypes.c_size_t, ctypes.c_int]
        L.oracle_sign_batch.argtypes = [ctypes.c_int, _u8p, _u8p, _u64p, _u64p, _u8p, ctypes.c_size_t,
                                        _i32p, ctypes.c_int]
        L.oracle_sign_rnd_batch.argtypes = [ctypes.c_int, _u8p, _u8p, _u64p, _u64p, _u8p, _u8p,
                                            ctypes.c_size_t, _i32p, ctypes.c_int]
        L.oracle_verify_batch.argtypes = [ctypes.c_int, _i32p, _u8p, _u8p, _u64p, _u64p, _u8p,
                                          ctypes.c_size_t, ctypes.c_int]
        L.oracle_max_threads.restype = ctypes.c_int
        L.oracle_elementwise.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _i32p,
                                         ctypes.c_size_t]
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


def params(mode: int) -> Params:
    p = Params()
    if lib().oracle_params(mode, ctypes.byref(p)) != 0:
        raise ValueError(f"unknown Dilithium mode {mode}")
    return p


def shake128(data: bytes, outlen: int) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
    out = np.zeros(outlen, dtype=np.uint8)
    lib().oracle_shake128(_p(out, _u8p), outlen, _p(buf, _u8p), len(data))
    return out.tobytes()


def shake256(data: bytes, outlen: int) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
    out = np.zeros(outlen, dtype=np.uint8)
    lib().oracle_shake256(_p(out, _u8p), outlen, _p(buf, _u8p), len(data))
    return out.tobytes()


def keccakf1600(state: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.uint64).copy()
    lib().oracle_keccakf1600(_p(s, _u64p))
    return s


def zetas() -> np.ndarray:
    ptr = lib().oracle_zetas()
    return np.ctypeslib.as_array(ptr, shape=(256,)).copy()


def ntt(a: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.int32).reshape(-1, 256).copy()
    lib().oracle_ntt_batch(_p(x, _i32p), x.shape[0], 0)
    return x.reshape(np.shape(a))


def invntt(a: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.int32).reshape(-1, 256).copy()
    lib().oracle_ntt_batch(_p(x, _i32p), x.shape[0], 1)
    return x.reshape(np.shape(a))


def ntt_pipeline(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    assert a.shape == b.shape and a.shape[-1] == 256
    c = np.empty_like(a)
    lib().oracle_ntt_pipeline(_p(c, _i32p), _p(a, _i32p), _p(b, _i32p), a.size // 256)
    return c


def keypair_batch(mode: int, seeds: np.ndarray, nthreads: int = 0):
    p = params(mode)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint8).reshape(-1, 32)
    n = seeds.shape[0]
    pk = np.zeros((n, p.pk_bytes), np.uint8)
    sk = np.zeros((n, p.sk_bytes), np.uint8)
    lib().oracle_keypair_batch(mode, _p(pk, _u8p), _p(sk, _u8p), _p(seeds, _u8p), n, nthreads)
    return pk, sk


def keypair_counter_batch(mode: int, n: int, rand_index: int = 0, nthreads: int = 0):
    p = params(mode)
    pk = np.zeros((n, p.pk_bytes), np.uint8)
    sk = np.zeros((n, p.sk_bytes), np.uint8)
    lib().oracle_keypair_counter_batch(mode, _p(pk, _u8p), _p(sk, _u8p), rand_index, n, nthreads)
    return pk, sk


def _msgs(msgs):
    """Accepts a 2-D uint8 array (fixed mlen) or a list of byte strings."""
    if isinstance(msgs, np.ndarray) and msgs.ndim == 2:
        n, ml = msgs.shape
        flat = np.ascontiguousarray(msgs, dtype=np.uint8).reshape(-1)
        off = np.arange(n, dtype=np.uint64) * ml
        ln = np.full(n, ml, dtype=np.uint64)
    else:
        lens = [len(m) for m in msgs]
        n = len(lens)
        flat = np.frombuffer(b"".join(bytes(m) for m in msgs), dtype=np.uint8).copy()
        off = np.zeros(n, dtype=np.uint64)
        if n > 1:
            off[1:] = np.cumsum(lens[:-1])
        ln = np.asarray(lens, dtype=np.uint64)
    if flat.size == 0:
        flat = np.zeros(1, np.uint8)
    return flat, off, ln, n


def sign_batch(mode: int, msgs, sk: np.ndarray, nthreads: int = 0):
    p = params(mode)
    flat, off, ln, n = _msgs(msgs)
    sk = np.ascontiguousarray(sk, dtype=np.uint8)
    sig = np.zeros((n, p.sig_bytes), np.uint8)
    att = np.zeros(n, np.int32)
    lib().oracle_sign_batch(mode, _p(sig, _u8p), _p(flat, _u8p), _p(off, _u64p), _p(ln, _u64p),
                            _p(sk, _u8p), n, _p(att, _i32p), nthreads)
    return sig, att


def sign_rnd_batch(mode: int, msgs, sk: np.ndarray, rhoprime: np.ndarray, nthreads: int = 0):
    """Randomized signatures with caller-chosen 64-byte rhoprime per item."""
    p = params(mode)
    flat, off, ln, n = _msgs(msgs)
    sk = np.ascontiguousarray(sk, dtype=np.uint8)
    rnd = np.ascontiguousarray(rhoprime, dtype=np.uint8).reshape(n, 64)
    sig = np.zeros((n, p.sig_bytes), np.uint8)
    att = np.zeros(n, np.int32)
    lib().oracle_sign_rnd_batch(mode, _p(sig, _u8p), _p(flat, _u8p), _p(off, _u64p), _p(ln, _u64p),
                                _p(sk, _u8p), _p(rnd, _u8p), n, _p(att, _i32p), nthreads)
    return sig, att


def verify_batch(mode: int, sig: np.ndarray, msgs, pk: np.ndarray, nthreads: int = 0):
    flat, off, ln, n = _msgs(msgs)
    sig = np.ascontiguousarray(sig, dtype=np.uint8)
    pk = np.ascontiguousarray(pk, dtype=np.uint8)
    ret = np.zeros(n, np.int32)
    lib().oracle_verify_batch(mode, _p(ret, _i32p), _p(sig, _u8p), _p(flat, _u8p), _p(off, _u64p),
                              _p(ln, _u64p), _p(pk, _u8p), n, nthreads)
    return ret


ELEMENTWISE_OPS = {"power2round": 0, "decompose": 1, "make_hint": 2, "use_hint": 3,
                   "reduce32": 4, "caddq": 5, "montgomery": 6}


def elementwise(op: str, a, b=None, mode: int = 2):
    """Returns (hi, lo) arrays of the named scalar routine applied elementwise."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(np.zeros_like(a) if b is None else b, dtype=np.int32)
    lo = np.zeros_like(a)
    hi = np.zeros_like(a)
    lib().oracle_elementwise(mode, ELEMENTWISE_OPS[op], _p(a, _i32p), _p(b, _i32p), _p(lo, _i32p),
                             _p(hi, _i32p), a.size)
    return hi, lo


def max_threads() -> int:
    return lib().oracle_max_threads()
