"""Pins two oracle functions against literal emulations of the reference's
own device code, run on the CPU with numpy:

1. The radix-8 NTT kernels (src/ntt.cu:58-207 with the load / store pattern
   of include/ntt.cuh:30-62): 32 lanes x 8 registers, shared-memory
   exchanges at strides 36 and 9, the zeta index of every butterfly, and the
   Montgomery product in its mulhi form (include/ntt.cuh:64-72).  The int32
   outputs must equal oracle.ntt / oracle.invntt word for word, which pins
   the representatives (not only the values mod q) of config 2.
2. The hint decoding of gpu_verify (src/verify.cu:52-93), whose failure
   masks are the XOR of per-iteration warp ballots.  For single errors it
   agrees with the specification form the oracle (and the engine) follow;
   two non-zero padding bytes 32 positions apart cancel in the reference
   and are accepted there, while the specification rejects them.  The test
   pins the oracle's verdict (-1) on that input and shows the emulated
   reference decode accepting it.
"""
import numpy as np
import pytest

import oracle

Q = 8380417
QINV = 58728449
LANES = 32


def _wrap32(v):
    return ((np.asarray(v, np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)).astype(np.int64)


def _mont(x, y):
    """mulhi(x, y) - mulhi(lo(x * y) * QINV, q) in 32-bit two's complement."""
    prod = np.asarray(x, np.int64) * np.asarray(y, np.int64)
    a_hi = prod >> 32
    a_lo = _wrap32(prod)
    t = _wrap32(a_lo * QINV)
    t_hi = (t * Q) >> 32
    return _wrap32(a_hi - t_hi)


class _Warp:
    """Registers regs[i] of the 32 lanes of a warp for P polynomials at once:
    array (8, P, 32); zetas indexed per lane."""

    def __init__(self, npoly):
        self.t = np.arange(LANES)
        self.smem = np.zeros((npoly, 256 + 32), np.int64)
        self.z = oracle.zetas().astype(np.int64)

    def zeta(self, idx):
        return self.z[np.asarray(idx)]  # scalar or per-lane (32,)

    def store(self, regs, addr):
        for i in range(8):
            self.smem[:, addr(i)] = regs[i]

    def load(self, regs, addr):
        for i in range(8):
            regs[i] = self.smem[:, addr(i)]


def _ntt_butt(regs, a, b, z):
    t = _mont(z, regs[b])
    regs[b] = _wrap32(regs[a] - t)
    regs[a] = _wrap32(regs[a] + t)


def _invntt_butt(regs, a, b, z):
    t = regs[a].copy()
    regs[a] = _wrap32(t + regs[b])
    regs[b] = _mont(z, _wrap32(t - regs[b]))


def ref_radix8_ntt(polys):
    p = np.asarray(polys, np.int64).reshape(-1, 256)
    w = _Warp(p.shape[0])
    t = w.t
    regs = np.zeros((8, p.shape[0], LANES), np.int64)
    for i in range(8):  # g_poly[i * 32 + threadIdx.x]
        regs[i] = p[:, i * 32 + t]
    z = w.zeta
    for a, b in ((0, 4), (1, 5), (2, 6), (3, 7)):
        _ntt_butt(regs, a, b, z(1))
    for a, b, k in ((0, 2, 2), (1, 3, 2), (4, 6, 3), (5, 7, 3)):
        _ntt_butt(regs, a, b, z(k))
    for a, b, k in ((0, 1, 4), (2, 3, 5), (4, 5, 6), (6, 7, 7)):
        _ntt_butt(regs, a, b, z(k))
    w.store(regs, lambda i: i * 36 + t)
    w.load(regs, lambda i: (t // 4) * 36 + (t & 3) + i * 4)
    for a, b in ((0, 4), (1, 5), (2, 6), (3, 7)):
        _ntt_butt(regs, a, b, z(8 + t // 4))
    for a, b, o in ((0, 2, 0), (1, 3, 0), (4, 6, 1), (5, 7, 1)):
        _ntt_butt(regs, a, b, z(16 + o + (t // 4) * 2))
    for a, b, o in ((0, 1, 0), (2, 3, 1), (4, 5, 2), (6, 7, 3)):
        _ntt_butt(regs, a, b, z(32 + o + (t // 4) * 4))

    def row_addr(i):
        off = (t & 3) + i * 4
        return (t // 4) * 36 + off // 8 + off
    w.store(regs, row_addr)
    w.load(regs, lambda i: t * 9 + i)
    for a, b, o in ((0, 2, 0), (1, 3, 0), (4, 6, 1), (5, 7, 1)):
        _ntt_butt(regs, a, b, z(64 + o + t * 2))
    for a, b, o in ((0, 1, 0), (2, 3, 1), (4, 5, 2), (6, 7, 3)):
        _ntt_butt(regs, a, b, z(128 + o + t * 4))
    out = np.zeros_like(p)
    for i in range(8):  # g_poly[threadIdx.x * 8 + i]
        out[:, t * 8 + i] = regs[i]
    return out.astype(np.int32)


def ref_radix8_invntt(polys):
    p = np.asarray(polys, np.int64).reshape(-1, 256)
    w = _Warp(p.shape[0])
    t = w.t
    regs = np.zeros((8, p.shape[0], LANES), np.int64)
    for i in range(8):  # g_poly[threadIdx.x * 8 + i]
        regs[i] = p[:, t * 8 + i]

    def nz(idx):
        return -w.zeta(idx)
    for a, b, o in ((0, 1, 1), (2, 3, 2), (4, 5, 3), (6, 7, 4)):
        _invntt_butt(regs, a, b, nz(256 - t * 4 - o))
    for a, b, o in ((0, 2, 1), (1, 3, 1), (4, 6, 2), (5, 7, 2)):
        _invntt_butt(regs, a, b, nz(128 - t * 2 - o))
    w.store(regs, lambda i: t * 9 + i)

    def row_addr(i):
        off = (t & 3) + i * 4
        return (t // 4) * 36 + off // 8 + off
    w.load(regs, row_addr)
    for a, b, o in ((0, 1, 1), (2, 3, 2), (4, 5, 3), (6, 7, 4)):
        _invntt_butt(regs, a, b, nz(64 - (t // 4) * 4 - o))
    for a, b, o in ((0, 2, 1), (1, 3, 1), (4, 6, 2), (5, 7, 2)):
        _invntt_butt(regs, a, b, nz(32 - (t // 4) * 2 - o))
    for a, b in ((0, 4), (1, 5), (2, 6), (3, 7)):
        _invntt_butt(regs, a, b, nz(16 - t // 4 - 1))
    w.store(regs, lambda i: (t // 4) * 36 + (t & 3) + i * 4)
    w.load(regs, lambda i: i * 36 + t)
    for a, b, k in ((0, 1, 7), (2, 3, 6), (4, 5, 5), (6, 7, 4)):
        _invntt_butt(regs, a, b, nz(k))
    for a, b, k in ((0, 2, 3), (1, 3, 3), (4, 6, 2), (5, 7, 2)):
        _invntt_butt(regs, a, b, nz(k))
    for a, b in ((0, 4), (1, 5), (2, 6), (3, 7)):
        _invntt_butt(regs, a, b, nz(1))
    for i in range(8):
        regs[i] = _mont(41978, regs[i])
    out = np.zeros_like(p)
    for i in range(8):  # g_poly[i * 32 + threadIdx.x]
        out[:, i * 32 + t] = regs[i]
    return out.astype(np.int32)


def test_radix8_emulation_is_the_transform():
    """Sanity of the emulation itself: values mod q are the negacyclic NTT
    (INTT(NTT(a)) = MONT * a mod q), so an indexing slip in the emulation
    cannot pass the equality tests below by accident."""
    a = np.random.default_rng(0).integers(-Q + 1, Q, (4, 256)).astype(np.int32)
    b = ref_radix8_invntt(ref_radix8_ntt(a)).astype(np.int64)
    assert np.array_equal(b % Q, a.astype(np.int64) * ((1 << 32) % Q) % Q)


def test_oracle_ntt_equals_reference_radix8_int32():
    rng = np.random.default_rng(11)
    a = rng.integers(-Q, Q, (64, 256)).astype(np.int32)  # config 2's range [-q, q)
    a[0] = -Q
    a[1] = Q - 1
    a[2] = 0
    assert np.array_equal(oracle.ntt(a), ref_radix8_ntt(a))


def test_oracle_invntt_equals_reference_radix8_int32():
    rng = np.random.default_rng(12)
    a = rng.integers(-Q, Q, (64, 256)).astype(np.int32)
    b = rng.integers(-Q, Q, (64, 256)).astype(np.int32)
    fa = oracle.ntt(a)
    # INTT inputs as config 2 makes them: Montgomery products in (-q, q),
    # and forward-NTT outputs (|x| < 9q) directly
    prod = _mont(fa, b).astype(np.int32)
    assert np.array_equal(oracle.invntt(prod), ref_radix8_invntt(prod))
    assert np.array_equal(oracle.invntt(fa), ref_radix8_invntt(fa))
    assert np.array_equal(oracle.ntt_pipeline(a, b), ref_radix8_invntt(prod))


# --------------------------------------------------------- hint decoding

def ref_hint_decode_fails(h, omega, k):
    """gpu_verify's hint checks (src/verify.cu:52-93) for one signature, with
    32 lanes: per k, lanes i = pos + t, pos + t + 32, ... < end each ballot
    'not strictly increasing', and the warp fails if the XOR of the ballots
    is non-zero; the same for non-zero padding bytes in [pos, omega)."""
    h = np.asarray(h, np.int64)
    pos = 0

    def xor_ballots(lo, hi, pred):
        mask = 0
        for base in range(lo, hi, LANES):
            ballot = 0
            for t in range(LANES):
                i = base + t
                if i < hi and pred(i):
                    ballot |= 1 << t
            mask ^= ballot
        return mask

    for kk in range(k):
        end = int(h[omega + kk])
        if end < pos or end > omega:
            return True
        if xor_ballots(pos, end, lambda i: i > pos and h[i] <= h[i - 1]):
            return True
        pos = end
    return xor_ballots(pos, omega, lambda i: h[i] != 0) != 0


def test_hint_padding_pair_32_apart():
    """Dilithium5 (omega 75): honest signatures with fewer than 43 hints leave
    two padding positions 32 apart; Dilithium2/3 signatures carry too many
    hints for their omega (80 / 55) to have such a pair."""
    mode, n = 5, 128
    p = oracle.params(mode)
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=31)
    msgs = [bytes([i]) * 40 for i in range(n)]
    sig, _ = oracle.sign_batch(mode, msgs, sk)
    hoff = p.sig_bytes - (p.omega + p.k)
    # a signature whose hint count leaves room for two padding bytes 32 apart
    cand = [i for i in range(n) if int(sig[i, hoff + p.omega + p.k - 1]) + 32 < p.omega]
    assert cand, "no signature with enough hint padding"
    i = cand[0]
    h = sig[i, hoff:].copy()
    assert not ref_hint_decode_fails(h, p.omega, p.k)
    assert oracle.verify_batch(mode, sig[i:i + 1], msgs[i:i + 1], pk[i:i + 1])[0] == 0

    pos = int(h[p.omega + p.k - 1])
    one = sig[i:i + 1].copy()
    one[0, hoff + pos] = 1  # a single non-zero padding byte: both forms reject
    assert ref_hint_decode_fails(one[0, hoff:], p.omega, p.k)
    assert oracle.verify_batch(mode, one, msgs[i:i + 1], pk[i:i + 1])[0] == -1

    two = one.copy()
    two[0, hoff + pos + 32] = 1  # same lane, next iteration: the ballots cancel
    assert not ref_hint_decode_fails(two[0, hoff:], p.omega, p.k)  # the reference accepts (ret 0)
    assert oracle.verify_batch(mode, two, msgs[i:i + 1], pk[i:i + 1])[0] == -1  # the specification rejects
