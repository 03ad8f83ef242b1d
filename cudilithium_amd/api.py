"""Python mirror of the reference's batched interface (include/api.cuh:27-63).

Same names, argument meaning, pool ownership and return conventions as
crypto_sign_keypair / crypto_sign_signature / crypto_sign_verify, over the
C-ABI of libcudilithium_amd.so.  Host buffers are numpy arrays (pinned
HostBuffer arrays work too); pools are caller-allocated pitched device
buffers sized with the reference's formulas (reference_pool).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceBuffer, HostBuffer, PitchedBuffer, stream_sync

c_void_p, c_size_t, c_int = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
u64p, u32p, i32p = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int32)

POOL_KEYPAIR, POOL_SIGN, POOL_SIGN_TEMP, POOL_VERIFY = 0, 1, 2, 3
ENGINE_PERSISTENT, ENGINE_PERSTAGE = 0, 1
# dil_sign_ctx_status bits
STATUS_ROUND_CAP, STATUS_NO_SECRET = 1, 2
# message hashing flag of dil_sign_ctx_set_message_hashing / dil_verify_device*_ex
HASH_LONG_MESSAGES = 1
# flag of dil_sign_ctx_set_claim: guided claims near the end of the queue
CLAIM_GUIDED = 0x10000


class SignLutElement(ctypes.Structure):  # api.cuh:31-34
    _fields_ = [("done_flag", ctypes.c_int), ("nonce", ctypes.c_uint16)]


class TaskLutC(ctypes.Structure):  # api.cuh:41-49 (sign_lut as an array)
    _fields_ = [("sign_lut", ctypes.POINTER(SignLutElement)), ("sign_lut_len", c_size_t),
                ("h_exec_lut", c_void_p), ("h_done_lut", c_void_p), ("h_copy_lut", c_void_p),
                ("d_exec_lut", c_void_p), ("d_done_lut", c_void_p), ("d_copy_lut", c_void_p)]


_registered = False


def _register():
    global _registered
    if _registered:
        return
    R = _lib.register
    R("dil_keypair_pool_row_bytes", c_size_t, [c_int])
    R("dil_sign_pool_row_bytes", c_size_t, [c_int, c_size_t])
    R("dil_sign_temp_row_bytes", c_size_t, [c_int])
    R("dil_verify_pool_row_bytes", c_size_t, [c_int, c_size_t])
    R("dil_reference_pool_row_bytes", c_size_t, [c_int, c_int, c_size_t])
    R("dil_crypto_sign_keypair", c_int, [c_int, c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, c_size_t])
    sig_args = [c_int, c_void_p, c_size_t, ctypes.POINTER(c_size_t), c_void_p, c_size_t, c_size_t, c_void_p,
                c_void_p, c_size_t, c_void_p, c_size_t, ctypes.POINTER(TaskLutC), c_size_t, c_size_t, c_void_p]
    R("dil_crypto_sign_signature", c_int, sig_args)
    R("dil_crypto_sign_signature_perstage", c_int, sig_args)
    R("dil_crypto_sign_verify", c_int, [c_int, c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, c_size_t, c_size_t,
                                        c_void_p, c_void_p, c_size_t, c_size_t, c_void_p])
    R("dil_keypair_work_row_bytes", c_size_t, [c_int])
    R("dil_keypair_device", c_int, [c_int, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_size_t,
                                    ctypes.c_uint64, c_size_t, c_void_p, c_size_t, c_void_p])
    R("dil_sign_ctx_create", c_int, [ctypes.POINTER(c_void_p), c_int, c_size_t, c_size_t])
    R("dil_sign_ctx_destroy", c_int, [c_void_p])
    R("dil_sign_ctx_set_tiling", c_int, [c_void_p, c_int, c_int, c_int])
    R("dil_sign_ctx_enable_timing", c_int, [c_void_p, c_int])
    R("dil_sign_ctx_timing", c_int, [c_void_p, c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)])
    R("dil_sign_ctx_timed_calls", c_int, [c_void_p])
    R("dil_sign_device", c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p,
                                 c_size_t, c_int, c_void_p])
    R("dil_sign_ctx_attempt_counts", c_int, [c_void_p, c_void_p, c_size_t, c_void_p])
    R("dil_sign_ctx_status", c_int, [c_void_p, ctypes.POINTER(c_int), c_void_p])
    R("dil_sign_ctx_executed_attempts", c_int, [c_void_p, ctypes.POINTER(ctypes.c_uint64), c_void_p])
    R("dil_verify_work_row_bytes", c_size_t, [c_int])
    R("dil_verify_device", c_int, [c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p,
                                   c_void_p, c_size_t, c_void_p, c_size_t, c_void_p])
    pp = ctypes.POINTER(c_void_p)
    R("dil_sign_graph_create", c_int, [pp, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p,
                                       c_void_p, c_size_t])
    R("dil_verify_graph_create", c_int, [pp, c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p,
                                         c_void_p, c_void_p, c_size_t, c_void_p, c_size_t])
    R("dil_keypair_graph_create", c_int, [pp, c_int, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_size_t,
                                          ctypes.c_uint64, c_size_t, c_void_p, c_size_t])
    R("dil_keystore_row_bytes", c_size_t, [c_int])
    R("dil_keystore_create", c_int, [pp, c_int, c_size_t])
    R("dil_keystore_destroy", c_int, [c_void_p])
    R("dil_keystore_capacity", c_size_t, [c_void_p])
    R("dil_keystore_mode", c_int, [c_void_p])
    R("dil_keystore_has_secret", c_int, [c_void_p, c_size_t])
    R("dil_release_host_staging", c_int, [])
    R("dil_set_kernel_tiling", c_int, [c_int, c_int, c_int, c_int])
    R("dil_get_kernel_tiling", c_int, [c_int, c_int, c_size_t])
    R("dil_sign_ctx_set_message_hashing", c_int, [c_void_p, c_int])
    R("dil_verify_device_ex", c_int, [c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p,
                                      c_void_p, c_size_t, c_void_p, c_size_t, c_int, c_void_p])
    R("dil_verify_device_resident_ex", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p,
                                               c_void_p, c_size_t, c_void_p, c_size_t, c_int, c_void_p])
    R("dil_keystore_load", c_int, [c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_size_t, c_size_t, c_void_p])
    R("dil_keystore_keygen", c_int, [c_void_p, c_size_t, c_void_p, c_size_t, ctypes.c_uint64, c_size_t, c_void_p])
    R("dil_keystore_export", c_int, [c_void_p, c_size_t, c_size_t, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p])
    R("dil_sign_ctx_create_resident", c_int, [pp, c_int, c_size_t, c_size_t])
    R("dil_sign_device_resident", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p,
                                          c_void_p, c_size_t, c_int, c_void_p])
    R("dil_verify_device_resident", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p,
                                            c_void_p, c_size_t, c_void_p, c_size_t, c_void_p])
    R("dil_sign_device_randomized", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t,
                                            c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_int,
                                            c_void_p])
    R("dil_sign_ctx_profile_phases", c_int, [c_void_p, c_int])
    R("dil_sign_ctx_set_claim", c_int, [c_void_p, c_int])
    R("dil_sign_ctx_phase_cycles", c_int, [c_void_p, c_void_p, c_void_p])
    R("dil_graph_launch", c_int, [c_void_p, c_void_p])
    R("dil_graph_batch", c_size_t, [c_void_p])
    R("dil_graph_destroy", c_int, [c_void_p])
    _registered = True


_tuned = False


def L():
    """The library, with its C signatures registered and the autotuner's
    recorded kernel tilings (tuning.json) applied on first use."""
    global _tuned
    _register()
    lib = _lib.load()
    if not _tuned:
        _tuned = True
        from . import autotune
        autotune.apply_cache(lib)
    return lib


def sizes(mode: int):
    lib = L()
    pk, sk, sig = lib.dil_public_key_bytes(mode), lib.dil_secret_key_bytes(mode), lib.dil_signature_bytes(mode)
    if not pk:
        raise ValueError(f"unknown Dilithium mode {mode}")
    return pk, sk, sig


def reference_row_bytes(mode: int, which: int, mlen: int = 0) -> int:
    return L().dil_reference_pool_row_bytes(mode, which, mlen)


def reference_pool(mode: int, which: int, rows: int, mlen: int = 0) -> PitchedBuffer:
    """cudaMallocPitch of `rows` rows sized by the reference's formula."""
    return PitchedBuffer(reference_row_bytes(mode, which, mlen), rows)


def _hp(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class TaskLut:
    """task_lut as allocated by the reference's callers
    (tests/test_vectors.cu:104-113): exec_threshold entries each."""

    def __init__(self, batch: int, exec_threshold: int):
        self.exec_threshold = exec_threshold
        self.sign_lut = (SignLutElement * max(1, batch))()
        self.h_exec = HostBuffer((exec_threshold,), np.uint32)
        self.h_done = HostBuffer((exec_threshold,), np.uint8)
        self.h_copy = HostBuffer((exec_threshold, 2), np.uint64)
        self.d_exec = DeviceBuffer(4 * exec_threshold)
        self.d_done = DeviceBuffer(exec_threshold)
        self.d_copy = DeviceBuffer(16 * exec_threshold)
        self.c = TaskLutC(ctypes.cast(self.sign_lut, ctypes.POINTER(SignLutElement)), batch, self.h_exec.ptr,
                          self.h_done.ptr, self.h_copy.ptr, self.d_exec.ptr, self.d_done.ptr, self.d_copy.ptr)


class TaskLutSlice:
    """Entries [first, first + count) of a TaskLut's exec / done / copy tables
    with a sign_lut of its own, as the reference's stream mode hands each
    stream a slice (benches/bench_dilithium.cu:205-217)."""

    def __init__(self, lut: TaskLut, first: int, count: int, batch: int):
        assert 0 < count and first + count <= lut.exec_threshold
        self.exec_threshold = count
        self.sign_lut = (SignLutElement * max(1, batch))()
        self.c = TaskLutC(ctypes.cast(self.sign_lut, ctypes.POINTER(SignLutElement)), batch,
                          lut.h_exec.ptr + 4 * first, lut.h_done.ptr + first, lut.h_copy.ptr + 16 * first,
                          lut.d_exec.ptr + 4 * first, lut.d_done.ptr + first, lut.d_copy.ptr + 16 * first)


def crypto_sign_signature_rows(mode, sig_ptr: int, sig_pitch: int, m_ptr: int | None, m_pitch: int, mlen: int,
                               sk_ptr: int, sign_pool_ptr: int, sign_pitch: int, temp_pool_ptr: int,
                               temp_pitch: int, lut, exec_threshold: int, batch_size: int, stream=None,
                               engine: int = ENGINE_PERSISTENT) -> int:
    """crypto_sign_signature on raw addresses (host rows and pool rows of a
    slice); `lut` is a TaskLut or TaskLutSlice.  Returns siglen."""
    fn = L().dil_crypto_sign_signature if engine == ENGINE_PERSISTENT else L().dil_crypto_sign_signature_perstage
    siglen = c_size_t(0)
    r = fn(mode, sig_ptr, sig_pitch, ctypes.byref(siglen), m_ptr, m_pitch, mlen, sk_ptr, sign_pool_ptr, sign_pitch,
           temp_pool_ptr, temp_pitch, ctypes.byref(lut.c), exec_threshold, batch_size, stream)
    _lib.check(r, "crypto_sign_signature")
    return siglen.value


def crypto_sign_keypair(mode, pk: np.ndarray, sk: np.ndarray, d_pool: PitchedBuffer, pitch: int,
                        batch_size: int = 1, stream=None, rand_index: int = 0) -> int:
    r = L().dil_crypto_sign_keypair(mode, _hp(pk), _hp(sk), d_pool.ptr, pitch, batch_size, stream, rand_index)
    _lib.check(r, "crypto_sign_keypair")
    return r


def crypto_sign_signature(mode, sig: np.ndarray, sig_pitch: int, m: np.ndarray, m_pitch: int, mlen: int,
                          sk: np.ndarray, d_sign_pool: PitchedBuffer, sign_pitch: int,
                          d_temp_pool: PitchedBuffer, temp_pitch: int, lut: TaskLut, exec_threshold: int,
                          batch_size: int = 1, stream=None, engine: int = ENGINE_PERSISTENT) -> int:
    """Returns siglen."""
    fn = L().dil_crypto_sign_signature if engine == ENGINE_PERSISTENT else L().dil_crypto_sign_signature_perstage
    siglen = c_size_t(0)
    mp = _hp(m) if m is not None and m.size else None
    r = fn(mode, _hp(sig), sig_pitch, ctypes.byref(siglen), mp, m_pitch, mlen, _hp(sk), d_sign_pool.ptr,
           sign_pitch, d_temp_pool.ptr, temp_pitch, ctypes.byref(lut.c), exec_threshold, batch_size, stream)
    _lib.check(r, "crypto_sign_signature")
    return siglen.value


def crypto_sign_verify(mode, ret: np.ndarray, sig: np.ndarray, sig_pitch: int, siglen: int, m: np.ndarray,
                       m_pitch: int, mlen: int, pk: np.ndarray, d_pool: PitchedBuffer, pitch: int,
                       batch_size: int = 1, stream=None) -> None:
    assert ret.dtype == np.int32
    mp = _hp(m) if m is not None and m.size else None
    r = L().dil_crypto_sign_verify(mode, _hp(ret), _hp(sig), sig_pitch, siglen, mp, m_pitch, mlen, _hp(pk),
                                   d_pool.ptr, pitch, batch_size, stream)
    _lib.check(r, "crypto_sign_verify")


class Dilithium:
    """Convenience wrapper over the drop-in calls for one mode: allocates
    pools sized by the reference's formulas and returns numpy arrays."""

    def __init__(self, mode: int):
        self.mode = mode
        self.pk_bytes, self.sk_bytes, self.sig_bytes = sizes(mode)

    def keypair(self, batch: int, rand_index: int = 0):
        pool = reference_pool(self.mode, POOL_KEYPAIR, batch)
        pk = np.zeros((batch, self.pk_bytes), np.uint8)
        sk = np.zeros((batch, self.sk_bytes), np.uint8)
        crypto_sign_keypair(self.mode, pk, sk, pool, pool.pitch, batch, None, rand_index)
        stream_sync(None)
        return pk, sk

    def sign(self, m: np.ndarray, sk: np.ndarray, exec_threshold: int = 2048, engine: int = ENGINE_PERSISTENT):
        m = np.ascontiguousarray(m, dtype=np.uint8)
        batch, mlen = m.shape
        pool = reference_pool(self.mode, POOL_SIGN, batch, mlen)
        temp = reference_pool(self.mode, POOL_SIGN_TEMP, exec_threshold)
        lut = TaskLut(batch, exec_threshold)
        sig = np.zeros((batch, self.sig_bytes), np.uint8)
        siglen = crypto_sign_signature(self.mode, sig, self.sig_bytes, m, mlen, mlen, np.ascontiguousarray(sk),
                                       pool, pool.pitch, temp, temp.pitch, lut, exec_threshold, batch, None, engine)
        stream_sync(None)
        assert siglen == self.sig_bytes
        return sig

    def verify(self, sig: np.ndarray, m: np.ndarray, pk: np.ndarray, siglen: int | None = None):
        m = np.ascontiguousarray(m, dtype=np.uint8)
        batch, mlen = m.shape
        pool = reference_pool(self.mode, POOL_VERIFY, batch, mlen)
        ret = np.zeros(batch, np.int32)
        crypto_sign_verify(self.mode, ret, np.ascontiguousarray(sig), self.sig_bytes,
                           self.sig_bytes if siglen is None else siglen, m, mlen, mlen, np.ascontiguousarray(pk),
                           pool, pool.pitch, batch, None)
        stream_sync(None)
        return ret


# ---------------------------------------------------------------- device API

def _pack_msgs(msgs):
    """List of byte strings -> (flat uint8, offsets u64, lengths u32) with
    8-byte aligned offsets."""
    lens = np.array([len(m) for m in msgs], dtype=np.uint32)
    offs = np.zeros(len(msgs), dtype=np.uint64)
    pos = 0
    for i, n in enumerate(lens):
        offs[i] = pos
        pos += (int(n) + 7) // 8 * 8
    flat = np.zeros(max(8, pos), dtype=np.uint8)
    for i, m in enumerate(msgs):
        if len(m):
            flat[int(offs[i]):int(offs[i]) + len(m)] = np.frombuffer(bytes(m), np.uint8)
    return flat, offs, lens


def keypair_device(mode: int, seeds: np.ndarray | None = None, n: int | None = None, rand_index: int = 0):
    pkb, skb, _ = sizes(mode)
    if seeds is not None:
        seeds = np.ascontiguousarray(seeds, dtype=np.uint8).reshape(-1, 32)
        n = seeds.shape[0]
    d_pk, d_sk = DeviceBuffer(n * pkb), DeviceBuffer(n * skb)
    wrow = (L().dil_keypair_work_row_bytes(mode) + 255) // 256 * 256
    d_work = DeviceBuffer(n * wrow)
    d_seeds = DeviceBuffer.from_array(seeds) if seeds is not None else None
    _lib.check(L().dil_keypair_device(mode, d_pk.ptr, pkb, d_sk.ptr, skb, d_seeds.ptr if d_seeds else None, 32,
                                      rand_index, n, d_work.ptr, wrow, None), "keypair_device")
    return d_pk.download((n, pkb), np.uint8), d_sk.download((n, skb), np.uint8)


def verify_device(mode: int, sig: np.ndarray, msgs, pk: np.ndarray):
    _, _, sigb = sizes(mode)
    n = len(msgs)
    flat, offs, lens = _pack_msgs(msgs)
    d_sig, d_pk = DeviceBuffer.from_array(np.ascontiguousarray(sig)), DeviceBuffer.from_array(np.ascontiguousarray(pk))
    d_msg, d_off, d_len = DeviceBuffer.from_array(flat), DeviceBuffer.from_array(offs), DeviceBuffer.from_array(lens)
    wrow = L().dil_verify_work_row_bytes(mode)
    d_work = DeviceBuffer(n * wrow)
    d_ret = DeviceBuffer(4 * n)
    _lib.check(L().dil_verify_device(mode, d_ret.ptr, d_sig.ptr, sig.shape[1], d_pk.ptr, pk.shape[1], d_msg.ptr,
                                     d_off.ptr, d_len.ptr, n, d_work.ptr, wrow, None), "verify_device")
    return d_ret.download((n,), np.int32)


class DeviceMessages:
    """A batch of messages resident in device memory (flat bytes plus per-item
    8-byte-aligned offsets and lengths)."""

    def __init__(self, msgs):
        flat, offs, lens = _pack_msgs(msgs)
        self.n = len(msgs)
        self.total_bytes = int(lens.sum())
        self.data, self.offs, self.lens = (DeviceBuffer.from_array(flat), DeviceBuffer.from_array(offs),
                                           DeviceBuffer.from_array(lens))


class SignContext:
    """dil_sign_ctx: per-item rows and attempt slots for batches of up to
    max_batch signatures of one mode."""

    def __init__(self, mode: int, max_batch: int, nslots: int | None = None, resident_only: bool = False):
        """nslots None: the autotuner's choice for this mode and batch bucket
        (cudilithium_amd/tuning.json), else one slot per item.  resident_only:
        rows without per-item expanded arrays (sign_resident only)."""
        from . import autotune
        tuned = autotune.lookup(mode, max_batch)
        if nslots is None:
            nslots = max(1, int(max_batch * (tuned[0] if tuned else 1.0)))
        self.mode, self.max_batch, self.nslots = mode, max_batch, nslots
        self.pk_bytes, self.sk_bytes, self.sig_bytes = sizes(mode)
        p = c_void_p()
        create = L().dil_sign_ctx_create_resident if resident_only else L().dil_sign_ctx_create
        _lib.check(create(ctypes.byref(p), mode, max_batch, nslots), "sign_ctx_create")
        self.handle = p.value
        if tuned and nslots // tuned[1] >= 1:
            self.set_tiling(tuned[1], tuned[2])
            claim = tuned[3] | (CLAIM_GUIDED if tuned[4] else 0)
            if claim:
                self.set_claim(claim)

    slots_per_block = 0
    occupancy = 0
    claim = 0

    def set_tiling(self, slots_per_block: int = 0, occupancy: int = 0, blocks: int = 0) -> None:
        _lib.check(L().dil_sign_ctx_set_tiling(self.handle, slots_per_block, occupancy, blocks), "set_tiling")
        self.slots_per_block, self.occupancy = slots_per_block, occupancy

    def set_message_hashing(self, flags: int) -> None:
        """0, or HASH_LONG_MESSAGES: mu / rhoprime with one sponge per 16-lane row."""
        _lib.check(L().dil_sign_ctx_set_message_hashing(self.handle, int(flags)), "set_message_hashing")

    def set_claim(self, items_per_block: int) -> None:
        """items per workgroup (0: its slots), optionally | CLAIM_GUIDED."""
        _lib.check(L().dil_sign_ctx_set_claim(self.handle, int(items_per_block)), "set_claim")
        self.claim = int(items_per_block)

    def enable_timing(self, max_calls: int) -> None:
        _lib.check(L().dil_sign_ctx_enable_timing(self.handle, int(max_calls)), "enable_timing")

    def timings(self):
        """[(fixed_ms, loop_ms)] of the timed calls since enable_timing."""
        out = []
        for i in range(L().dil_sign_ctx_timed_calls(self.handle)):
            f, l = ctypes.c_float(), ctypes.c_float()
            _lib.check(L().dil_sign_ctx_timing(self.handle, i, ctypes.byref(f), ctypes.byref(l)), "timing")
            out.append((f.value, l.value))
        return out

    def attempt_counts(self, n: int) -> np.ndarray:
        out = np.zeros(max(1, n), np.uint32)
        _lib.check(L().dil_sign_ctx_attempt_counts(self.handle, out.ctypes.data, n, None), "attempt_counts")
        return out[:n]

    def sign_device(self, d_sig: DeviceBuffer, d_sk: DeviceBuffer, msgs: DeviceMessages,
                    engine: int = ENGINE_PERSISTENT, stream=None) -> None:
        _lib.check(L().dil_sign_device(self.handle, d_sig.ptr, self.sig_bytes, d_sk.ptr, self.sk_bytes,
                                       msgs.data.ptr, msgs.offs.ptr, msgs.lens.ptr, msgs.n, engine, stream),
                   "sign_device")

    def sign_resident_device(self, store: "KeyStore", d_key_idx: DeviceBuffer, d_sig: DeviceBuffer,
                             msgs: DeviceMessages, engine: int = ENGINE_PERSISTENT, stream=None) -> None:
        _lib.check(L().dil_sign_device_resident(self.handle, store.handle, d_key_idx.ptr, d_sig.ptr, self.sig_bytes,
                                                msgs.data.ptr, msgs.offs.ptr, msgs.lens.ptr, msgs.n, engine, stream),
                   "sign_device_resident")

    def sign_resident(self, store: "KeyStore", key_idx, msgs, engine: int = ENGINE_PERSISTENT) -> np.ndarray:
        n = len(msgs)
        dm = DeviceMessages(msgs)
        d_idx = DeviceBuffer.from_array(np.ascontiguousarray(key_idx, dtype=np.uint32))
        d_sig = DeviceBuffer(n * self.sig_bytes)
        self.sign_resident_device(store, d_idx, d_sig, dm, engine)
        st = self.status()
        if st & STATUS_ROUND_CAP:
            raise _lib.DilError("persistent engine stopped at its round limit")
        if st & STATUS_NO_SECRET:
            raise _lib.DilError("a key of the batch has no secret key (verify-only)")
        return d_sig.download((n, self.sig_bytes), np.uint8)

    def sign_randomized(self, msgs, rhoprime: np.ndarray, sk: np.ndarray | None = None,
                        store: "KeyStore | None" = None, key_idx=None, engine: int = ENGINE_PERSISTENT) -> np.ndarray:
        """Randomized signatures: rhoprime[i] (64 bytes) for item i, under host
        secret keys `sk` or resident keys `key_idx` of `store`."""
        n = len(msgs)
        dm = DeviceMessages(msgs)
        d_rnd = DeviceBuffer.from_array(np.ascontiguousarray(rhoprime, dtype=np.uint8).reshape(n, 64))
        d_sk = DeviceBuffer.from_array(np.ascontiguousarray(sk, dtype=np.uint8)) if sk is not None else None
        d_idx = DeviceBuffer.from_array(np.ascontiguousarray(key_idx, dtype=np.uint32)) if store else None
        d_sig = DeviceBuffer(n * self.sig_bytes)
        _lib.check(L().dil_sign_device_randomized(self.handle, store.handle if store else None,
                                                  d_idx.ptr if d_idx else None, d_sig.ptr, self.sig_bytes,
                                                  d_sk.ptr if d_sk else None, self.sk_bytes, dm.data.ptr, dm.offs.ptr,
                                                  dm.lens.ptr, d_rnd.ptr, 64, n, engine, None), "sign_randomized")
        st = self.status()
        if st & STATUS_ROUND_CAP:
            raise _lib.DilError("persistent engine stopped at its round limit")
        if st & STATUS_NO_SECRET:
            raise _lib.DilError("a key of the batch has no secret key (verify-only)")
        return d_sig.download((n, self.sig_bytes), np.uint8)

    def graph(self, d_sig: DeviceBuffer, d_sk: DeviceBuffer, msgs: DeviceMessages) -> "Graph":
        """The persistent-engine call sign_device(d_sig, d_sk, msgs) captured
        as a HIP graph (replayed by Graph.launch)."""
        p = c_void_p()
        _lib.check(L().dil_sign_graph_create(ctypes.byref(p), self.handle, d_sig.ptr, self.sig_bytes, d_sk.ptr,
                                             self.sk_bytes, msgs.data.ptr, msgs.offs.ptr, msgs.lens.ptr, msgs.n),
                   "sign_graph_create")
        return Graph(p.value, keep=(self, d_sig, d_sk, msgs))

    PHASES = ("assign", "y", "w", "c", "z", "commit")

    def profile_phases(self, enable: bool) -> None:
        _lib.check(L().dil_sign_ctx_profile_phases(self.handle, 1 if enable else 0), "profile_phases")

    def phase_cycles(self) -> dict:
        out = np.zeros(6, np.uint64)
        _lib.check(L().dil_sign_ctx_phase_cycles(self.handle, out.ctypes.data, None), "phase_cycles")
        return dict(zip(self.PHASES, (int(v) for v in out)))

    def executed_attempts(self) -> int:
        v = ctypes.c_uint64(0)
        _lib.check(L().dil_sign_ctx_executed_attempts(self.handle, ctypes.byref(v), None), "executed_attempts")
        return v.value

    def status(self) -> int:
        s = c_int(0)
        _lib.check(L().dil_sign_ctx_status(self.handle, ctypes.byref(s), None), "status")
        return s.value

    def sign(self, msgs, sk: np.ndarray, engine: int = ENGINE_PERSISTENT) -> np.ndarray:
        n = len(msgs)
        dm = DeviceMessages(msgs)
        d_sk = DeviceBuffer.from_array(np.ascontiguousarray(sk, dtype=np.uint8))
        d_sig = DeviceBuffer(n * self.sig_bytes)
        self.sign_device(d_sig, d_sk, dm, engine)
        if engine == ENGINE_PERSISTENT and self.status() != 0:
            raise _lib.DilError("persistent engine stopped at its round limit")
        return d_sig.download((n, self.sig_bytes), np.uint8)

    def __del__(self):
        try:
            if self.handle:
                L().dil_sign_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Graph:
    """dil_graph: one captured keygen / sign / verify call.  Holds references
    to the buffers it was captured with so they outlive it."""

    def __init__(self, handle: int, keep=()):
        self.handle, self._keep = handle, keep
        self.n = L().dil_graph_batch(handle)

    def launch(self, stream=None) -> None:
        _lib.check(L().dil_graph_launch(self.handle, stream), "graph_launch")

    def __del__(self):
        try:
            if self.handle:
                L().dil_graph_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def verify_graph(mode: int, d_ret: DeviceBuffer, d_sig: DeviceBuffer, d_pk: DeviceBuffer, msgs: DeviceMessages,
                 d_work: DeviceBuffer, work_pitch: int) -> Graph:
    """dil_verify_device over these buffers captured as a HIP graph."""
    pkb, _, sigb = sizes(mode)
    p = c_void_p()
    _lib.check(L().dil_verify_graph_create(ctypes.byref(p), mode, d_ret.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb,
                                           msgs.data.ptr, msgs.offs.ptr, msgs.lens.ptr, msgs.n, d_work.ptr,
                                           work_pitch), "verify_graph_create")
    return Graph(p.value, keep=(d_ret, d_sig, d_pk, msgs, d_work))


def keypair_graph(mode: int, d_pk: DeviceBuffer, d_sk: DeviceBuffer, n: int, d_work: DeviceBuffer, work_pitch: int,
                  d_seeds: DeviceBuffer | None = None, rand_index: int = 0) -> Graph:
    """dil_keypair_device over these buffers captured as a HIP graph."""
    pkb, skb, _ = sizes(mode)
    p = c_void_p()
    _lib.check(L().dil_keypair_graph_create(ctypes.byref(p), mode, d_pk.ptr, pkb, d_sk.ptr, skb,
                                            d_seeds.ptr if d_seeds else None, 32, rand_index, n, d_work.ptr,
                                            work_pitch), "keypair_graph_create")
    return Graph(p.value, keep=(d_pk, d_sk, d_work, d_seeds))


class KeyStore:
    """dil_keystore: `capacity` resident keys of one mode with their expanded
    A, s1^, s2^, t0^, t1^ and tr."""

    def __init__(self, mode: int, capacity: int):
        p = c_void_p()
        _lib.check(L().dil_keystore_create(ctypes.byref(p), mode, capacity), "keystore_create")
        self.handle, self.mode, self.capacity = p.value, mode, capacity
        self.pk_bytes, self.sk_bytes, self.sig_bytes = sizes(mode)

    def load(self, first: int, pk: np.ndarray, sk: np.ndarray | None = None) -> None:
        """Install host keys at [first, first + len(pk)) and expand them."""
        pk = np.ascontiguousarray(pk, dtype=np.uint8)
        skp = np.ascontiguousarray(sk, dtype=np.uint8) if sk is not None else None
        _lib.check(L().dil_keystore_load(self.handle, first, pk.ctypes.data, self.pk_bytes,
                                         skp.ctypes.data if skp is not None else None, self.sk_bytes, pk.shape[0],
                                         None), "keystore_load")
        stream_sync(None)

    def keygen(self, first: int, seeds: np.ndarray | None = None, n: int | None = None, rand_index: int = 0) -> None:
        d_seeds = None
        if seeds is not None:
            seeds = np.ascontiguousarray(seeds, dtype=np.uint8).reshape(-1, 32)
            n = seeds.shape[0]
            d_seeds = DeviceBuffer.from_array(seeds)
        _lib.check(L().dil_keystore_keygen(self.handle, first, d_seeds.ptr if d_seeds else None, 32, rand_index, n,
                                           None), "keystore_keygen")
        stream_sync(None)

    def export(self, first: int, n: int, with_sk: bool = True):
        pk = np.zeros((n, self.pk_bytes), np.uint8)
        sk = np.zeros((n, self.sk_bytes), np.uint8) if with_sk else None
        _lib.check(L().dil_keystore_export(self.handle, first, n, pk.ctypes.data, self.pk_bytes,
                                           sk.ctypes.data if with_sk else None, self.sk_bytes, None),
                   "keystore_export")
        stream_sync(None)
        return pk, sk

    def has_secret(self, idx: int) -> bool:
        return bool(L().dil_keystore_has_secret(self.handle, idx))

    def verify(self, key_idx, sig: np.ndarray, msgs, flags: int | None = None) -> np.ndarray:
        """Verdicts under resident keys; flags (e.g. HASH_LONG_MESSAGES) go
        to dil_verify_device_resident_ex, None to dil_verify_device_resident."""
        n = len(msgs)
        dm = DeviceMessages(msgs)
        d_idx = DeviceBuffer.from_array(np.ascontiguousarray(key_idx, dtype=np.uint32))
        d_sig = DeviceBuffer.from_array(np.ascontiguousarray(sig, dtype=np.uint8))
        wrow = L().dil_verify_work_row_bytes(self.mode)
        d_work, d_ret = DeviceBuffer(n * wrow), DeviceBuffer(4 * n)
        if flags is None:
            _lib.check(L().dil_verify_device_resident(self.handle, d_idx.ptr, d_ret.ptr, d_sig.ptr, sig.shape[1],
                                                      dm.data.ptr, dm.offs.ptr, dm.lens.ptr, n, d_work.ptr, wrow,
                                                      None), "verify_device_resident")
        else:
            _lib.check(L().dil_verify_device_resident_ex(self.handle, d_idx.ptr, d_ret.ptr, d_sig.ptr, sig.shape[1],
                                                         dm.data.ptr, dm.offs.ptr, dm.lens.ptr, n, d_work.ptr, wrow,
                                                         int(flags), None), "verify_device_resident_ex")
        return d_ret.download((n,), np.int32)

    def __del__(self):
        try:
            if self.handle:
                L().dil_keystore_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ------------------------------------------------------------------ serving

OP_KEYGEN, OP_SIGN, OP_VERIFY = 0, 1, 2


class ServerConfig(ctypes.Structure):  # dil_server_config
    _fields_ = [("hbm_budget_bytes", c_size_t), ("lanes", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("max_keys", ctypes.c_uint32 * 3), ("batch_window_us", ctypes.c_uint32),
                ("long_message_bytes", ctypes.c_uint32), ("long_lanes", ctypes.c_uint32)]


class Request(ctypes.Structure):  # dil_request
    _fields_ = [("op", ctypes.c_int32), ("mode", ctypes.c_int32), ("key", ctypes.c_uint32), ("mlen", ctypes.c_uint32),
                ("msg", c_void_p), ("seed", c_void_p), ("sig", c_void_p), ("out", c_void_p),
                ("status", ctypes.c_int32), ("verdict", ctypes.c_int32),
                ("t_submit_ns", ctypes.c_uint64), ("t_launch_ns", ctypes.c_uint64), ("t_done_ns", ctypes.c_uint64),
                ("batch_id", ctypes.c_uint32), ("done", ctypes.c_int32)]


class ServerStats(ctypes.Structure):  # dil_server_stats
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("items_by_op", ctypes.c_uint64 * 3), ("lane_bytes", ctypes.c_uint64),
                ("max_msg_bytes", ctypes.c_uint64), ("batches_by_op", ctypes.c_uint64 * 3),
                ("stage_ns", ctypes.c_uint64 * 3), ("device_ns", ctypes.c_uint64 * 3),
                ("finish_ns", ctypes.c_uint64 * 3)]


def _register_server():
    R = _lib.register
    pp = ctypes.POINTER(c_void_p)
    R("dil_server_create", c_int, [pp, ctypes.POINTER(ServerConfig)])
    R("dil_server_destroy", c_int, [c_void_p])
    R("dil_server_keystore", c_void_p, [c_void_p, c_int])
    R("dil_server_submit", c_int, [c_void_p, c_void_p, c_size_t])
    R("dil_server_wait", c_int, [c_void_p, c_void_p, ctypes.c_int64])
    R("dil_server_drain", c_int, [c_void_p])
    R("dil_server_get_stats", c_int, [c_void_p, ctypes.POINTER(ServerStats)])
    R("dil_server_replay", c_int, [c_void_p, c_void_p, c_void_p, c_size_t])
    R("dil_monotonic_ns", ctypes.c_uint64, [])


class _StoreView(KeyStore):
    """A server's key store (owned by the server, not freed here)."""

    def __init__(self, handle, mode, capacity):  # noqa: super().__init__ allocates; this wraps
        self.handle, self.mode, self.capacity = handle, mode, capacity
        self.pk_bytes, self.sk_bytes, self.sig_bytes = sizes(mode)

    def __del__(self):
        pass


class Server:
    """dil_server: continuous batching of keygen / sign / verify requests with
    resident keys.  Requests are `Request` structs whose buffers the caller
    keeps alive until done (Server.requests builds such arrays)."""

    def __init__(self, hbm_budget_bytes: int, lanes: int = 1, max_batch: int = 4096, max_keys=(0, 0, 0),
                 batch_window_us: int = 0, long_message_bytes: int = 4096, long_lanes: int = 0):
        L()
        _register_server()
        cfg = ServerConfig(hbm_budget_bytes, lanes, max_batch, (ctypes.c_uint32 * 3)(*max_keys), batch_window_us,
                           long_message_bytes, long_lanes)
        p = c_void_p()
        _lib.check(L().dil_server_create(ctypes.byref(p), ctypes.byref(cfg)), "server_create")
        self.handle = p.value
        self.stores = {}
        for i, m in enumerate((2, 3, 5)):
            if max_keys[i]:
                self.stores[m] = _StoreView(L().dil_server_keystore(self.handle, m), m, max_keys[i])

    def submit(self, reqs) -> None:
        """reqs: a ctypes array of Request (all submitted at once)."""
        ptrs = (c_void_p * len(reqs))(*[ctypes.addressof(r) for r in reqs])
        _lib.check(L().dil_server_submit(self.handle, ptrs, len(reqs)), "server_submit")

    def replay(self, reqs, arrival_ns: np.ndarray) -> None:
        arr = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        _lib.check(L().dil_server_replay(self.handle, ctypes.addressof(reqs), arr.ctypes.data, len(reqs)),
                   "server_replay")

    def drain(self) -> None:
        _lib.check(L().dil_server_drain(self.handle), "server_drain")

    def wait(self, req, timeout_us: int = -1) -> int:
        return L().dil_server_wait(self.handle, ctypes.addressof(req), timeout_us)

    def stats(self) -> ServerStats:
        st = ServerStats()
        _lib.check(L().dil_server_get_stats(self.handle, ctypes.byref(st)), "server_stats")
        return st

    def close(self) -> None:
        if getattr(self, "handle", None):
            L().dil_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RequestBatch:
    """A ctypes array of n Requests plus the host buffers they point to:
    messages (concatenated), signatures to verify, seeds, and outputs."""

    def __init__(self, ops, modes, keys, msgs, sigs=None, seeds=None):
        n = len(ops)
        self.n = n
        self.reqs = (Request * n)()
        self._msgs = [np.frombuffer(bytes(m), np.uint8) if len(m) else np.zeros(1, np.uint8) for m in msgs]
        out_bytes = [0] * n
        self.out = []
        self._sigs, self._seeds = [], []
        self._sigs_by_index = {}
        for i in range(n):
            op, mode = int(ops[i]), int(modes[i])
            pkb, skb, sigb = sizes(mode)
            r = self.reqs[i]
            r.op, r.mode, r.key = op, mode, int(keys[i])
            r.mlen = len(msgs[i])
            r.msg = self._msgs[i].ctypes.data
            if op == OP_KEYGEN:
                sd = np.ascontiguousarray(seeds[i], dtype=np.uint8)
                self._seeds.append(sd)
                r.seed = sd.ctypes.data
                o = np.zeros(pkb + skb, np.uint8)
            elif op == OP_SIGN:
                o = np.zeros(sigb, np.uint8)
            else:
                sg = np.ascontiguousarray(sigs[i], dtype=np.uint8)
                self._sigs.append(sg)
                self._sigs_by_index[i] = sg
                r.sig = sg.ctypes.data
                o = np.zeros(1, np.uint8)
            self.out.append(o)
            r.out = o.ctypes.data

    def field(self, name) -> np.ndarray:
        return np.array([getattr(r, name) for r in self.reqs])
