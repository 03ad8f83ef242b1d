"""Message hashing with one sponge per 16-lane row (DIL_HASH_LONG_MESSAGES):
signatures and verdicts equal the oracle's at message lengths around every
SHAKE256 block boundary of tr || m (32 + mlen a multiple of 136 and one byte
either side), up to 64 KiB, for host and resident keys, deterministic and
randomized signing, and with the default one-lane hashing on the same inputs."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _lengths():
    ls = {0, 1, 3, 4, 5, 7, 8, 9, 65535, 65536}
    for k in (1, 2, 3, 30, 31, 481):
        b = 136 * k - 32
        ls |= {b - 1, b, b + 1}
    return sorted(ls)


def _msgs(seed, lens):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, n, dtype=np.uint8).tobytes() for n in lens]


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_row_hashing_sign_matches_oracle(mode):
    from cudilithium_amd import api
    lens = _lengths()
    n = len(lens)
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=70 + mode)
    msgs = _msgs(mode, lens)
    exp, exp_att = oracle.sign_batch(mode, msgs, sk)
    for engine in (api.ENGINE_PERSISTENT, api.ENGINE_PERSTAGE):
        ctx = api.SignContext(mode, n, n)
        ctx.set_message_hashing(api.HASH_LONG_MESSAGES)
        assert np.array_equal(ctx.sign(msgs, sk, engine), exp), engine
    # resident keys and randomized signing through the same hashing
    ks = api.KeyStore(mode, n)
    ks.load(0, pk, sk)
    idx = np.arange(n, dtype=np.uint32)[::-1].copy()
    exp_r, _ = oracle.sign_batch(mode, msgs, sk[idx])
    ctx = api.SignContext(mode, n, n, resident_only=True)
    ctx.set_message_hashing(api.HASH_LONG_MESSAGES)
    assert np.array_equal(ctx.sign_resident(ks, idx, msgs), exp_r)
    rnd = np.random.default_rng(9).integers(0, 256, (n, 64), dtype=np.uint8)
    exp_rnd, _ = oracle.sign_rnd_batch(mode, msgs, sk, rnd)
    ctx = api.SignContext(mode, n, n)
    ctx.set_message_hashing(api.HASH_LONG_MESSAGES)
    assert np.array_equal(ctx.sign_randomized(msgs, rnd, sk=sk), exp_rnd)


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_row_hashing_verify_matches_oracle(mode):
    from cudilithium_amd import api
    from cudilithium_amd.device import DeviceBuffer, synchronize
    lens = _lengths()
    n = len(lens)
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=170 + mode)
    msgs = _msgs(100 + mode, lens)
    sig, _ = oracle.sign_batch(mode, msgs, sk)
    bad = np.arange(1, n, 3)
    for i in bad:
        sig[i, (i * 37) % sig.shape[1]] ^= 1 << (i % 8)
    want = oracle.verify_batch(mode, sig, msgs, pk)
    assert (want[bad] == -1).all()
    pkb, _, sigb = api.sizes(mode)
    L = api.L()
    d_sig, d_pk, dm = DeviceBuffer.from_array(sig), DeviceBuffer.from_array(pk), api.DeviceMessages(msgs)
    wrow = L.dil_verify_work_row_bytes(mode)
    d_work, d_ret = DeviceBuffer(n * wrow), DeviceBuffer(4 * n)
    for flags in (api.HASH_LONG_MESSAGES, 0):
        d_ret.memset(0x33)
        api._lib.check(L.dil_verify_device_ex(mode, d_ret.ptr, d_sig.ptr, sigb, d_pk.ptr, pkb, dm.data.ptr,
                                              dm.offs.ptr, dm.lens.ptr, n, d_work.ptr, wrow, flags, None), "verify")
        synchronize()
        assert np.array_equal(d_ret.download((n,), np.int32), want), flags
    # resident keys, an index past the store rejected
    ks = api.KeyStore(mode, n)
    ks.load(0, pk)
    idx = np.arange(n, dtype=np.uint32)
    idx[2] = n + 5
    d_idx = DeviceBuffer.from_array(idx)
    want_r = want.copy()
    want_r[2] = -1
    d_ret.memset(0x33)
    api._lib.check(L.dil_verify_device_resident_ex(ks.handle, d_idx.ptr, d_ret.ptr, d_sig.ptr, sigb, dm.data.ptr,
                                                   dm.offs.ptr, dm.lens.ptr, n, d_work.ptr, wrow,
                                                   api.HASH_LONG_MESSAGES, None), "verify_resident")
    synchronize()
    assert np.array_equal(d_ret.download((n,), np.int32), want_r)
