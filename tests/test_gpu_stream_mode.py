"""The reference's stream mode (benches/bench_dilithium.cu:180-249) against
the drop-in C-ABI: several host threads, each with its own stream and
contiguous slices of the batch, of the pitched pools and of the task tables,
call keypair / sign / verify concurrently, with pinned and with pageable host
buffers.  Every byte and verdict must equal the oracle's (and hence the
single-call batch form's)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("host", ["pinned", "pageable"])
@pytest.mark.parametrize("mode", [2, 3])
def test_stream_mode_threads_on_slices_match_oracle(mode, host):
    from cudilithium_amd import api
    from cudilithium_amd.device import HostBuffer, Stream
    n, mlen, thr, ns = 600, 45, 256, 4
    pkb, skb, sigb = api.sizes(mode)
    kpool = api.reference_pool(mode, api.POOL_KEYPAIR, n)
    spool = api.reference_pool(mode, api.POOL_SIGN, n, mlen)
    tpool = api.reference_pool(mode, api.POOL_SIGN_TEMP, thr)
    vpool = api.reference_pool(mode, api.POOL_VERIFY, n, mlen)
    lut = api.TaskLut(n, thr)
    rng = np.random.default_rng(mode * 10 + (host == "pinned"))
    msg = rng.integers(0, 256, (n, mlen), dtype=np.uint8)
    if host == "pinned":
        bufs = [HostBuffer((n, pkb), np.uint8), HostBuffer((n, skb), np.uint8), HostBuffer((n, mlen), np.uint8),
                HostBuffer((n, sigb), np.uint8), HostBuffer((n,), np.int32)]
        pk, sk, m, sig, ret = (b.array for b in bufs)
    else:
        pk, sk, m = np.zeros((n, pkb), np.uint8), np.zeros((n, skb), np.uint8), np.zeros((n, mlen), np.uint8)
        sig, ret = np.zeros((n, sigb), np.uint8), np.zeros(n, np.int32)
    m[...] = msg
    ret[...] = 7
    streams = [Stream() for _ in range(ns)]
    per, tper = n // ns, thr // ns
    L = api.L()

    def work(i):
        f, c = i * per, per + (n % ns if i == ns - 1 else 0)
        tf, tc = i * tper, tper + (thr % ns if i == ns - 1 else 0)
        st = streams[i].handle
        # rand_index 2 f: item j of the slice gets counter seed 2 (f + j) + 1,
        # the same key as item f + j of one whole-batch call
        api._lib.check(L.dil_crypto_sign_keypair(mode, pk.ctypes.data + f * pkb, sk.ctypes.data + f * skb,
                                                 kpool.ptr + f * kpool.pitch, kpool.pitch, c, st, 2 * f), "keypair")
        streams[i].synchronize()
        api.crypto_sign_signature_rows(mode, sig.ctypes.data + f * sigb, sigb, m.ctypes.data + f * mlen, mlen, mlen,
                                       sk.ctypes.data + f * skb, spool.ptr + f * spool.pitch, spool.pitch,
                                       tpool.ptr + tf * tpool.pitch, tpool.pitch,
                                       api.TaskLutSlice(lut, tf, tc, c), tc, c, st)
        api._lib.check(L.dil_crypto_sign_verify(mode, ret.ctypes.data + 4 * f, sig.ctypes.data + f * sigb, sigb,
                                                sigb, m.ctypes.data + f * mlen, mlen, mlen, pk.ctypes.data + f * pkb,
                                                vpool.ptr + f * vpool.pitch, vpool.pitch, c, st), "verify")
        streams[i].synchronize()

    with ThreadPoolExecutor(max_workers=ns) as ex:
        list(ex.map(work, range(ns)))
    epk, esk = oracle.keypair_counter_batch(mode, n, rand_index=0)
    assert np.array_equal(pk, epk) and np.array_equal(sk, esk)
    esig, _ = oracle.sign_batch(mode, msg, esk)
    assert np.array_equal(sig, esig)
    assert (ret == 0).all()
