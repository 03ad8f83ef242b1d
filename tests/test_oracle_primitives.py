"""Pins the oracle's building blocks against independent definitions:
hashlib (SHAKE), FIPS 202 known answers, closed forms and the reference's own
table entries (zetas, src/ntt.cu:24-56), O(N^2) polynomial evaluation (NTT),
and exhaustive checks of the rounding routines against their defining
equations (include/rounding.cuh, include/reduce.cuh)."""
import hashlib

import numpy as np
import pytest

import oracle

Q = 8380417
MONT = (1 << 32) % Q


def brv8(i):
    return int(f"{i:08b}"[::-1], 2)


@pytest.mark.parametrize("inlen", [0, 1, 7, 8, 31, 32, 33, 66, 135, 136, 137, 167, 168, 169, 271, 272,
                                   273, 500, 1312, 4096 + 32])
def test_shake_matches_hashlib(inlen):
    rng = np.random.default_rng(inlen)
    data = rng.integers(0, 256, inlen, dtype=np.uint8).tobytes()
    for outlen in (1, 32, 64, 136, 168, 680, 840, 1001):
        assert oracle.shake128(data, outlen) == hashlib.shake_128(data).digest(outlen)
        assert oracle.shake256(data, outlen) == hashlib.shake_256(data).digest(outlen)


def test_keccak_zero_state_known_answer():
    # FIPS 202 / Keccak team intermediate values: Keccak-f[1600] of the all-zero
    # state, lane (0,0) and lane (1,0).
    s = oracle.keccakf1600(np.zeros(25, np.uint64))
    assert int(s[0]) == 0xF1258F7940E1DDE7
    assert int(s[1]) == 0x84D5CCF933C0478A


def test_zetas_closed_form_and_reference_entries():
    z = oracle.zetas()
    assert z[0] == 0
    for i in range(1, 256):
        v = MONT * pow(1753, brv8(i), Q) % Q
        if v > (Q - 1) // 2:
            v -= Q
        assert z[i] == v
    # entries as printed in the reference table, src/ntt.cu:25 and :56
    assert list(z[:8]) == [0, 25847, -2608894, -518909, 237124, -777960, -876248, 466468]
    assert list(z[-8:]) == [-554416, 3919660, -48306, -1362209, 3937738, 1400424, -846154, 1976782]


def test_ntt_is_evaluation_at_odd_powers():
    """NTT(a)[i] = a(1753^(2*brv8(i)+1)) mod q, checked by O(N^2) evaluation."""
    rng = np.random.default_rng(1)
    a = rng.integers(-Q + 1, Q, (3, 256)).astype(np.int32)
    out = oracle.ntt(a).astype(np.int64)
    xs = np.array([pow(1753, 2 * brv8(i) + 1, Q) for i in range(256)], dtype=np.int64)
    powers = np.ones((256, 256), dtype=np.int64)
    for j in range(1, 256):
        powers[:, j] = powers[:, j - 1] * xs % Q
    for p in range(3):
        ev = np.zeros(256, dtype=np.int64)
        for j in range(256):
            ev = (ev + (int(a[p, j]) % Q) * powers[:, j]) % Q
        assert np.array_equal(out[p] % Q, ev)


def test_invntt_inverts_with_montgomery_factor():
    rng = np.random.default_rng(2)
    a = rng.integers(-Q + 1, Q, (16, 256)).astype(np.int32)
    b = oracle.invntt(oracle.ntt(a))
    assert np.array_equal(b.astype(np.int64) % Q, a.astype(np.int64) * MONT % Q)
    assert np.all(np.abs(b) < Q)
    one = np.ones((1, 256), np.int32)
    r = oracle.invntt(oracle.ntt(one))
    # bench_ntt_radix2.cu:1263-1282: INTT(NTT(1)) is MONT (or MONT + q)
    assert set(np.unique(r).tolist()) <= {MONT - Q, MONT}


def test_ntt_pipeline_is_negacyclic_product():
    """INTT(NTT(a) o_mont NTT(b)) = a*b mod (X^256+1) (no extra factor)."""
    rng = np.random.default_rng(3)
    a = rng.integers(-Q + 1, Q, (2, 256)).astype(np.int32)
    b = rng.integers(-Q + 1, Q, (2, 256)).astype(np.int32)
    c = oracle.ntt_pipeline(a, oracle.ntt(b))
    for p in range(2):
        ref = np.zeros(512, dtype=object)
        for i in range(256):
            ref[i:i + 256] += int(a[p, i]) * b[p].astype(object)
        neg = (ref[:256] - ref[256:]) % Q
        assert np.array_equal(c[p].astype(np.int64) % Q, neg.astype(np.int64))


def test_montgomery_exact_identity():
    rng = np.random.default_rng(4)
    x = rng.integers(-(1 << 31), 1 << 31, 200000, dtype=np.int64)
    y = rng.integers(-Q, Q, 200000, dtype=np.int64)
    # keep |x*y| <= 2^31 * q as the routine requires
    x = np.clip(x, -(1 << 31) + 1, (1 << 31) - 1)
    r, _ = oracle.elementwise("montgomery", x.astype(np.int32), y.astype(np.int32))
    rinv = pow(1 << 32, -1, Q)
    exp = np.array([(int(a) * int(b) * rinv) % Q for a, b in zip(x[:2000], y[:2000])])
    assert np.array_equal(r[:2000].astype(np.int64) % Q, exp)
    assert np.all(np.abs(r) < Q)


def test_reduce32_and_caddq_exhaustive_ranges():
    a = np.arange(-(1 << 31), (1 << 31) - (1 << 22), 4099, dtype=np.int64).astype(np.int32)
    r, _ = oracle.elementwise("reduce32", a)
    assert np.array_equal(r.astype(np.int64) % Q, a.astype(np.int64) % Q)
    assert r.min() >= -6283009 and r.max() <= 6283007
    c, _ = oracle.elementwise("caddq", r)
    assert np.array_equal(c, np.where(r < 0, r + Q, r))


ALL = np.arange(Q, dtype=np.int32)


def test_power2round_exhaustive():
    a1, a0 = oracle.elementwise("power2round", ALL)
    assert np.array_equal(a1.astype(np.int64) * 8192 + a0, ALL)
    assert a0.min() > -4096 and a0.max() <= 4096


@pytest.mark.parametrize("mode,g2", [(2, (Q - 1) // 88), (3, (Q - 1) // 32)])
def test_decompose_exhaustive_matches_spec(mode, g2):
    a1, a0 = oracle.elementwise("decompose", ALL, mode=mode)
    r = ALL.astype(np.int64)
    r0 = r % (2 * g2)
    r0 = np.where(r0 > g2, r0 - 2 * g2, r0)
    wrap = (r - r0) == Q - 1
    e1 = np.where(wrap, 0, (r - r0) // (2 * g2))
    e0 = np.where(wrap, r0 - 1, r0)
    assert np.array_equal(a1, e1)
    assert np.array_equal(a0, e0)


@pytest.mark.parametrize("mode,g2,beta", [(2, (Q - 1) // 88, 78), (3, (Q - 1) // 32, 196),
                                          (5, (Q - 1) // 32, 120)])
def test_hint_recovers_high_bits(mode, g2, beta):
    """Correctness of MakeHint/UseHint as the protocol uses them: the signer
    hints w1 from (w0 - cs2 + ct0) and the verifier recovers w1 from
    w - cs2 + ct0 (mod q)."""
    rng = np.random.default_rng(mode)
    n = 2_000_000
    w = rng.integers(0, Q, n).astype(np.int32)
    w1, w0 = oracle.elementwise("decompose", w, mode=mode)
    e1 = rng.integers(-beta, beta + 1, n)
    e2 = rng.integers(-g2 + 1, g2, n)
    ok = np.abs(w0.astype(np.int64) - e1) < g2 - beta
    a0 = (w0.astype(np.int64) - e1 + e2).astype(np.int32)
    h, _ = oracle.elementwise("make_hint", a0, w1, mode=mode)
    wp = ((w.astype(np.int64) - e1 + e2) % Q).astype(np.int32)
    rec, _ = oracle.elementwise("use_hint", wp, h, mode=mode)
    assert np.array_equal(rec[ok], w1[ok])
    assert ok.mean() > 0.9
