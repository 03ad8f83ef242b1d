"""Batched NTT through the C-ABI (reference: include/ntt.cuh:30-62)."""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceBuffer


def forward_device(buf: DeviceBuffer, npoly: int, pitch_words: int = 256, stream=None) -> None:
    _lib.check(_lib.load().dil_ntt_forward(buf.ptr, npoly, pitch_words, stream), "ntt_forward")


def inverse_device(buf: DeviceBuffer, npoly: int, pitch_words: int = 256, stream=None) -> None:
    _lib.check(_lib.load().dil_ntt_inverse(buf.ptr, npoly, pitch_words, stream), "ntt_inverse")


def pipeline_device(c: DeviceBuffer, a: DeviceBuffer, b: DeviceBuffer, npoly: int, pitch_words: int = 256,
                    stream=None) -> None:
    _lib.check(_lib.load().dil_ntt_pointwise_pipeline(c.ptr, a.ptr, b.ptr, npoly, pitch_words, stream),
               "ntt_pipeline")


def forward(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    buf = DeviceBuffer.from_array(a)
    forward_device(buf, a.size // 256)
    return buf.download(a.shape, np.int32)


def inverse(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    buf = DeviceBuffer.from_array(a)
    inverse_device(buf, a.size // 256)
    return buf.download(a.shape, np.int32)


def pipeline(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    da, db = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b)
    dc = DeviceBuffer(a.nbytes)
    pipeline_device(dc, da, db, a.size // 256)
    return dc.download(a.shape, np.int32)


def keccak_f1600(states: np.ndarray) -> np.ndarray:
    """states: (n, 25) uint64; returns the permuted states."""
    s = np.ascontiguousarray(np.asarray(states, dtype=np.uint64).T)
    buf = DeviceBuffer.from_array(s)
    _lib.check(_lib.load().dil_keccak_f1600_batch(buf.ptr, s.shape[1], None), "keccak")
    return buf.download(s.shape, np.uint64).T.copy()
