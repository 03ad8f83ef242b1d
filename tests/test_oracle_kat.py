"""Pins the CPU oracle end to end: regenerating the reference's 10,000-vector
tvecs files for modes 2/3/5 must reproduce the SHA-256 values hard-coded in the
reference's tests/CMakeLists.txt:21-23."""
import numpy as np
import pytest

import oracle
from tests import kat


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_oracle_reproduces_reference_tvecs_hash(mode):
    m = kat.messages()
    pk, sk = oracle.keypair_counter_batch(mode, kat.NVECTORS, rand_index=0)
    sig, att = oracle.sign_batch(mode, m, sk)
    ret = oracle.verify_batch(mode, sig, m, pk)
    assert (ret == 0).all()
    assert kat.tvecs_sha256(m, pk, sk, sig, ret) == kat.PINS[mode]
    # expected repetitions of the round-3 specification: 4.25 / 5.1 / 3.85
    assert abs(att.mean() - {2: 4.25, 3: 5.1, 5: 3.85}[mode]) < 0.25


def test_planted_mistake_breaks_the_pin():
    """A one-byte change in one of 10,000 signatures must change the hash."""
    mode = 2
    n = kat.NVECTORS
    m = kat.messages()
    pk, sk = oracle.keypair_counter_batch(mode, n)
    sig, _ = oracle.sign_batch(mode, m, sk)
    sig[1234, 100] ^= 1
    ret = np.zeros(n, np.int32)
    assert kat.tvecs_sha256(m, pk, sk, sig, ret) != kat.PINS[mode]
