"""Resident key stores (dil_keystore_*): keygen into the store, loading host
keys, and signing / verifying under resident keys give the same bytes and
verdicts as the oracle (and hence as the non-resident entry points)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _msgs(seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, int(rng.integers(lo, hi + 1)), dtype=np.uint8).tobytes() for _ in range(n)]


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_keystore_keygen_sign_verify_match_oracle(mode):
    from cudilithium_amd import api
    nkeys, n = 24, 120
    seeds = np.random.default_rng(500 + mode).integers(0, 256, (nkeys, 32), dtype=np.uint8)
    epk, esk = oracle.keypair_batch(mode, seeds)
    ks = api.KeyStore(mode, nkeys)
    ks.keygen(0, seeds)
    pk, sk = ks.export(0, nkeys)
    assert np.array_equal(pk, epk) and np.array_equal(sk, esk)

    rng = np.random.default_rng(mode)
    key_idx = rng.integers(0, nkeys, n).astype(np.uint32)
    msgs = _msgs(600 + mode, n, 0, 900)
    exp, _ = oracle.sign_batch(mode, msgs, esk[key_idx])
    for engine in (api.ENGINE_PERSISTENT, api.ENGINE_PERSTAGE):
        ctx = api.SignContext(mode, n, 64 if engine == api.ENGINE_PERSTAGE else n, resident_only=True)
        got = ctx.sign_resident(ks, key_idx, msgs, engine)
        assert np.array_equal(got, exp), engine

    sig = exp.copy()
    flip = np.arange(3, n, 5)
    for i in flip:
        b = int(rng.integers(0, sig.shape[1] * 8))
        sig[i, b // 8] ^= 1 << (b % 8)
    ret = ks.verify(key_idx, sig, msgs)
    oret = oracle.verify_batch(mode, sig, msgs, epk[key_idx])
    assert np.array_equal(ret, oret)
    assert (ret[flip] == -1).all() and (np.delete(ret, flip) == 0).all()


def test_keystore_load_host_keys_and_verify_only_keys():
    from cudilithium_amd import api
    mode, nkeys, n = 3, 10, 40
    pk, sk = oracle.keypair_counter_batch(mode, nkeys, rand_index=900)
    full = api.KeyStore(mode, nkeys)
    full.load(0, pk[:6], sk[:6])
    full.load(6, pk[6:], sk[6:])
    key_idx = (np.arange(n) * 7 % nkeys).astype(np.uint32)
    msgs = _msgs(11, n, 1, 3000)
    exp, _ = oracle.sign_batch(mode, msgs, sk[key_idx])
    ctx = api.SignContext(mode, n, n)
    assert np.array_equal(ctx.sign_resident(full, key_idx, msgs), exp)
    # the non-resident path on the same context still works
    assert np.array_equal(ctx.sign(msgs, sk[key_idx]), exp)

    vonly = api.KeyStore(mode, nkeys)
    vonly.load(0, pk)  # public keys only
    ret = vonly.verify(key_idx, exp, msgs)
    assert (ret == 0).all()
    # wrong key for every item, and an index past the store's capacity
    bad_idx = ((key_idx + 1) % nkeys).astype(np.uint32)
    bad_idx[5] = nkeys + 3
    assert (vonly.verify(bad_idx, exp, msgs) == -1).all()


def test_verify_only_load_clears_secrets_and_signing_reports_it():
    """A verify-only load over rows that held signing keys leaves no secret
    behind: export of sk is refused, signing under such a key gives an
    all-zero signature with DIL_SIGN_STATUS_NO_SECRET (both engines), and the
    other items of the batch are still the oracle's signatures."""
    from cudilithium_amd import api, _lib
    mode, nkeys, n = 3, 6, 30
    pk, sk = oracle.keypair_counter_batch(mode, nkeys, rand_index=1200)
    ks = api.KeyStore(mode, nkeys)
    ks.load(0, pk, sk)
    assert all(ks.has_secret(i) for i in range(nkeys))
    other_pk, _ = oracle.keypair_counter_batch(mode, 2, rand_index=1300)
    ks.load(2, other_pk)  # keys 2 and 3 become verify-only
    assert [ks.has_secret(i) for i in range(nkeys)] == [True, True, False, False, True, True]
    with pytest.raises(_lib.DilError):
        ks.export(0, nkeys)
    pk_out, _ = ks.export(0, nkeys, with_sk=False)
    assert np.array_equal(pk_out[2:4], other_pk) and np.array_equal(pk_out[4:], pk[4:])
    assert np.array_equal(ks.export(4, 2)[1], sk[4:])

    key_idx = (np.arange(n) % nkeys).astype(np.uint32)
    msgs = _msgs(1400, n, 1, 500)
    good = np.nonzero(key_idx != 2)[0]
    good = good[key_idx[good] != 3]
    exp, _ = oracle.sign_batch(mode, [msgs[i] for i in good], sk[key_idx[good]])
    for engine in (api.ENGINE_PERSISTENT, api.ENGINE_PERSTAGE):
        ctx = api.SignContext(mode, n, n, resident_only=True)
        d_idx = api.DeviceBuffer.from_array(key_idx)
        dm = api.DeviceMessages(msgs)
        d_sig = api.DeviceBuffer(n * ctx.sig_bytes)
        d_sig.memset(0xAB)
        ctx.sign_resident_device(ks, d_idx, d_sig, dm, engine)
        assert ctx.status() == api.STATUS_NO_SECRET, engine
        got = d_sig.download((n, ctx.sig_bytes), np.uint8)
        bad = np.isin(key_idx, [2, 3])
        assert not got[bad].any(), engine
        assert np.array_equal(got[good], exp), engine
        with pytest.raises(_lib.DilError):
            ctx.sign_resident(ks, key_idx, msgs, engine)
        # a batch of signing keys only clears the status again
        assert np.array_equal(ctx.sign_resident(ks, key_idx[good], [msgs[i] for i in good], engine), exp)
        assert ctx.status() == 0
    # verification under the verify-only keys still works
    vsig, _ = oracle.sign_batch(mode, [b"m2", b"m3"], oracle.keypair_counter_batch(mode, 2, rand_index=1300)[1])
    assert (ks.verify(np.array([2, 3], np.uint32), vsig, [b"m2", b"m3"]) == 0).all()


def test_resident_only_context_rejects_plain_sign():
    from cudilithium_amd import api, _lib
    mode = 2
    ctx = api.SignContext(mode, 4, 4, resident_only=True)
    _, sk = oracle.keypair_counter_batch(mode, 4)
    with pytest.raises(_lib.DilError):
        ctx.sign([b"x"] * 4, sk)


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_randomized_signing_matches_oracle(mode):
    """dil_sign_device_randomized with caller rhoprime: host keys and resident
    keys, both engines, equal the oracle's randomized signatures."""
    from cudilithium_amd import api
    n, nkeys = 60, 6
    pk, sk = oracle.keypair_counter_batch(mode, nkeys, rand_index=300)
    key_idx = (np.arange(n) % nkeys).astype(np.uint32)
    msgs = _msgs(900 + mode, n, 0, 2000)
    rnd = np.random.default_rng(mode).integers(0, 256, (n, 64), dtype=np.uint8)
    exp, _ = oracle.sign_rnd_batch(mode, msgs, sk[key_idx], rnd)
    ctx = api.SignContext(mode, n, n)
    assert np.array_equal(ctx.sign_randomized(msgs, rnd, sk=sk[key_idx]), exp)
    assert np.array_equal(ctx.sign_randomized(msgs, rnd, sk=sk[key_idx], engine=api.ENGINE_PERSTAGE), exp)
    ks = api.KeyStore(mode, nkeys)
    ks.load(0, pk, sk)
    assert np.array_equal(ctx.sign_randomized(msgs, rnd, store=ks, key_idx=key_idx), exp)
    assert (oracle.verify_batch(mode, exp, msgs, pk[key_idx]) == 0).all()


def test_resident_verify_large_batch_forked_hashing():
    """Resident-key verification of more than 8,192 items, where tr and mu
    are hashed with two lanes per item on an auxiliary stream beside the
    main kernels (a partial last workgroup: 8,300 items): verdicts equal the
    oracle's, flipped items -1, through the plain and the long-message
    entry points."""
    from cudilithium_amd import api
    mode, nkeys, n = 2, 16, 8300
    seeds = np.random.default_rng(77).integers(0, 256, (nkeys, 32), dtype=np.uint8)
    epk, esk = oracle.keypair_batch(mode, seeds)
    ks = api.KeyStore(mode, nkeys)
    ks.keygen(0, seeds)
    rng = np.random.default_rng(78)
    key_idx = rng.integers(0, nkeys, n).astype(np.uint32)
    msgs = _msgs(79, n, 0, 300)
    ctx = api.SignContext(mode, n, resident_only=True)
    sig = ctx.sign_resident(ks, key_idx, msgs, api.ENGINE_PERSISTENT)
    flip = np.arange(5, n, 10)
    for i in flip:
        b = int(rng.integers(0, sig.shape[1] * 8))
        sig[i, b // 8] ^= 1 << (b % 8)
    oret = oracle.verify_batch(mode, sig, msgs, epk[key_idx])
    assert (oret[flip] == -1).all() and (np.delete(oret, flip) == 0).all()
    assert np.array_equal(ks.verify(key_idx, sig, msgs), oret)
    assert np.array_equal(ks.verify(key_idx, sig, msgs, flags=api.HASH_LONG_MESSAGES), oret)
