"""Randomized signing in the oracle (api.cu:143-148: rhoprime drawn at random
instead of SHAKE256(key || mu)).  Pinned by: (1) supplying the deterministic
rhoprime explicitly reproduces the KAT-pinned deterministic signature, and
(2) signatures under random rhoprime verify under the pinned verifier and
differ from the deterministic ones."""
import hashlib

import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_randomized_signing_pins(mode):
    pk, sk = oracle.keypair_counter_batch(mode, 12, rand_index=5)
    msgs = [bytes([i]) * (i * 37) for i in range(12)]
    det, det_att = oracle.sign_batch(mode, msgs, sk)
    rp = []
    for i, m in enumerate(msgs):
        mu = hashlib.shake_256(sk[i, 64:96].tobytes() + m).digest(64)
        rp.append(np.frombuffer(hashlib.shake_256(sk[i, 32:64].tobytes() + mu).digest(64), np.uint8))
    same, same_att = oracle.sign_rnd_batch(mode, msgs, sk, np.stack(rp))
    assert np.array_equal(same, det) and np.array_equal(same_att, det_att)
    rnd = np.random.default_rng(mode).integers(0, 256, (12, 64), dtype=np.uint8)
    sig, _ = oracle.sign_rnd_batch(mode, msgs, sk, rnd)
    assert (oracle.verify_batch(mode, sig, msgs, pk) == 0).all()
    assert all(not np.array_equal(sig[i], det[i]) for i in range(12))
