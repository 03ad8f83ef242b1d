"""Drop-in calls with pinned host buffers (hipHostMalloc, the reference's
cudaMallocHost in tests/test_vectors.cu), which the library's copy kernel
reads and writes directly, against pageable buffers (numpy), which go through
its pinned staging buffer; the C++ drop-in test pins the pageable path to the
reference's KAT hashes. Every copy of api.cu:55-58, 117-126, 443-446 and
504-513 is covered, with 16-, 4- and 1-byte access widths, an odd host pitch,
a pointer that starts inside its allocation, a pinned buffer rewritten by the
CPU between calls, and pageable copies over several 16 MiB staging chunks
(the staging buffer's two halves in turn)."""
import numpy as np
import pytest

from cudilithium_amd import api
from cudilithium_amd.device import HostBuffer

pytestmark = pytest.mark.gpu


def _pinned(a: np.ndarray, shape=None) -> HostBuffer:
    h = HostBuffer(shape or a.shape, a.dtype)
    h.array[...] = 0
    if a.size:
        h.array[tuple(slice(0, s) for s in a.shape)] = a
    return h


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_dropin_pinned_host_buffers_match_pageable(mode):
    n, mlen, thr = 67, 33, 128
    pkb, skb, sigb = api.sizes(mode)
    kp = api.reference_pool(mode, api.POOL_KEYPAIR, n)
    pk0, sk0 = np.zeros((n, pkb), np.uint8), np.zeros((n, skb), np.uint8)
    api.crypto_sign_keypair(mode, pk0, sk0, kp, kp.pitch, n, None, 5)
    api.stream_sync(None)
    hpk, hsk = HostBuffer((n, pkb), np.uint8), HostBuffer((n, skb), np.uint8)
    api.crypto_sign_keypair(mode, hpk.array, hsk.array, kp, kp.pitch, n, None, 5)
    api.stream_sync(None)
    assert np.array_equal(hpk.array, pk0) and np.array_equal(hsk.array, sk0)

    m = np.random.default_rng(mode).integers(0, 256, (n, mlen), dtype=np.uint8)
    sp = api.reference_pool(mode, api.POOL_SIGN, n, mlen)
    tp = api.reference_pool(mode, api.POOL_SIGN_TEMP, thr)
    lut = api.TaskLut(n, thr)
    sig0 = np.zeros((n, sigb), np.uint8)
    api.crypto_sign_signature(mode, sig0, sigb, m, mlen, mlen, sk0, sp, sp.pitch, tp, tp.pitch, lut, thr, n)
    api.stream_sync(None)
    hm = HostBuffer((n + 1, mlen), np.uint8)  # rows start one row into the allocation
    hm.array[1:] = m
    spitch = sigb + 5
    hsig = HostBuffer((n, spitch), np.uint8)
    hsig.array[...] = 0xA5
    api.crypto_sign_signature(mode, hsig.array, spitch, hm.array[1:], mlen, mlen, hsk.array, sp, sp.pitch, tp,
                              tp.pitch, lut, thr, n)
    api.stream_sync(None)
    assert np.array_equal(hsig.array[:, :sigb], sig0)
    assert (hsig.array[:, sigb:] == 0xA5).all()  # bytes between rows untouched
    # the CPU rewrites a pinned buffer the GPU has read: the next call sees the new bytes
    m2 = m[::-1].copy()
    sig2 = np.zeros((n, sigb), np.uint8)
    api.crypto_sign_signature(mode, sig2, sigb, m2, mlen, mlen, sk0, sp, sp.pitch, tp, tp.pitch, lut, thr, n)
    api.stream_sync(None)
    assert not np.array_equal(sig2, sig0)
    hm.array[1:] = m2
    api.crypto_sign_signature(mode, hsig.array, spitch, hm.array[1:], mlen, mlen, hsk.array, sp, sp.pitch, tp,
                              tp.pitch, lut, thr, n)
    api.stream_sync(None)
    assert np.array_equal(hsig.array[:, :sigb], sig2)
    hm.array[1:] = m

    sig0[3, 10] ^= 1
    vp = api.reference_pool(mode, api.POOL_VERIFY, n, mlen)
    ret0 = np.full(n, 7, np.int32)
    api.crypto_sign_verify(mode, ret0, sig0, sigb, sigb, m, mlen, mlen, pk0, vp, vp.pitch, n)
    api.stream_sync(None)
    want = np.zeros(n, np.int32)
    want[3] = -1
    assert np.array_equal(ret0, want)
    hsig2 = _pinned(sig0)
    hret = HostBuffer((n,), np.int32)
    hret.array[...] = 7
    api.crypto_sign_verify(mode, hret.array, hsig2.array, sigb, sigb, hm.array[1:], mlen, mlen, hpk.array, vp,
                           vp.pitch, n)
    api.stream_sync(None)
    assert np.array_equal(hret.array, want)


def test_dropin_pageable_copies_span_staging_chunks():
    mode, n, mlen, thr = 3, 8500, 16, 2048  # sk rows 4000 B: 34 MB, three 16 MiB chunks
    pkb, skb, sigb = api.sizes(mode)
    kp = api.reference_pool(mode, api.POOL_KEYPAIR, n)
    pk0, sk0 = np.zeros((n, pkb), np.uint8), np.zeros((n, skb), np.uint8)
    api.crypto_sign_keypair(mode, pk0, sk0, kp, kp.pitch, n, None, 11)
    api.stream_sync(None)
    hpk, hsk = HostBuffer((n, pkb), np.uint8), HostBuffer((n, skb), np.uint8)
    api.crypto_sign_keypair(mode, hpk.array, hsk.array, kp, kp.pitch, n, None, 11)
    api.stream_sync(None)
    assert np.array_equal(hpk.array, pk0) and np.array_equal(hsk.array, sk0)
    m = np.random.default_rng(9).integers(0, 256, (n, mlen), dtype=np.uint8)
    sp = api.reference_pool(mode, api.POOL_SIGN, n, mlen)
    tp = api.reference_pool(mode, api.POOL_SIGN_TEMP, thr)
    lut = api.TaskLut(n, thr)
    sig0 = np.zeros((n, sigb), np.uint8)
    api.crypto_sign_signature(mode, sig0, sigb, m, mlen, mlen, sk0, sp, sp.pitch, tp, tp.pitch, lut, thr, n)
    api.stream_sync(None)
    hm, hsig = _pinned(m), HostBuffer((n, sigb), np.uint8)
    api.crypto_sign_signature(mode, hsig.array, sigb, hm.array, mlen, mlen, hsk.array, sp, sp.pitch, tp, tp.pitch,
                              lut, thr, n)
    api.stream_sync(None)
    assert np.array_equal(hsig.array, sig0)
    # the staging buffer can be released between calls; the next pageable copy
    # allocates it again and gives the same bytes
    assert api.L().dil_release_host_staging() == 0
    assert api.L().dil_release_host_staging() == 0
    sig1 = np.zeros((n, sigb), np.uint8)
    api.crypto_sign_signature(mode, sig1, sigb, m, mlen, mlen, sk0, sp, sp.pitch, tp, tp.pitch, lut, thr, n)
    api.stream_sync(None)
    assert np.array_equal(sig1, sig0)
