"""Algorithmic work of the hot path, counted with fixed constants that do not
depend on the implementation (the figures roofline.achieved uses).

    OPS_KECCAK = 6240   32-bit ops per Keccak-f[1600]: 24 rounds x 130 64-bit
                        ops (theta 55, rho 24, chi 50, iota 1) x 2 halves
    OPS_BFLY   = 7      per NTT butterfly (Montgomery product 4 mul + 1 sub,
                        then one add and one sub)
    OPS_MONT   = 5      per pointwise Montgomery product or INTT scaling

Keccak permutation counts follow the specification's squeeze policy:
ExpandA 5 SHAKE128 blocks, ExpandMask 5 SHAKE256 blocks, ExpandS 2 blocks,
and a hash of inlen bytes costs floor(inlen / rate) + out_blocks.
"""
from __future__ import annotations

OPS_KECCAK = 6240
OPS_BFLY = 7
OPS_MONT = 5
N = 256
FWD_NTT_OPS = 1024 * OPS_BFLY
INV_NTT_OPS = 1024 * OPS_BFLY + N * OPS_MONT
POINTWISE_OPS = N * OPS_MONT
NTT_PIPELINE_OPS = FWD_NTT_OPS + POINTWISE_OPS + INV_NTT_OPS  # 16,896 per polynomial

PARAMS = {  # mode: (K, L, polyw1 bytes, pk bytes)
    2: (4, 4, 192, 1312),
    3: (6, 5, 128, 1952),
    5: (8, 7, 128, 2592),
}


def _hash_perms(inlen: int, rate: int = 136, out_blocks: int = 1) -> int:
    return inlen // rate + out_blocks


def sign_attempt_counts(mode: int):
    """(Keccak perms, forward transforms, inverse transforms, pointwise polys)
    of one signing attempt (sign.cu:463-1196)."""
    k, l, w1b, _ = PARAMS[mode]
    perms = 5 * l + _hash_perms(64 + k * w1b) + 1  # y, c~ = H(mu || w1), SampleInBall
    fwd = l + 1  # NTT(y), NTT(c)
    inv = k + l + 2 * k  # w, c s1, c s2, c t0
    pw = k * l + l + 2 * k
    return perms, fwd, inv, pw


def _ops(perms, fwd, inv, pw):
    return perms * OPS_KECCAK + fwd * FWD_NTT_OPS + inv * INV_NTT_OPS + pw * POINTWISE_OPS


def sign_attempt_ops(mode: int) -> int:
    return _ops(*sign_attempt_counts(mode))


def sign_fixed_ops(mode: int, mlen: int) -> int:
    """mu, rhoprime, ExpandA, NTT of s1, s2, t0 (api.cu:117-182)."""
    k, l, _, _ = PARAMS[mode]
    perms = _hash_perms(32 + mlen) + 1 + 5 * k * l
    return _ops(perms, l + 2 * k, 0, 0)


def verify_ops(mode: int, mlen: int) -> int:
    """tr, mu, ExpandA, SampleInBall, c~'; NTT of z, c, t1; A z; c t1; INTT
    (verify.cu:28-325)."""
    k, l, w1b, pkb = PARAMS[mode]
    perms = _hash_perms(pkb) + _hash_perms(32 + mlen) + 5 * k * l + 1 + _hash_perms(64 + k * w1b)
    return _ops(perms, l + 1 + k, k, k * l + k)


def verify_kernel_ops(mode: int, mlen: int) -> dict:
    """verify_ops split by the engine's three launches: vf_prep (tr, mu,
    SampleInBall), vf_main (ExpandA, NTT of z / c / t1, A z - c t1, INTT) and
    vf_final (c~' = H(mu || w1'))."""
    k, l, w1b, pkb = PARAMS[mode]
    prep = _ops(_hash_perms(pkb) + _hash_perms(32 + mlen) + 1, 0, 0, 0)
    main = _ops(5 * k * l, l + 1 + k, k, k * l + k)
    final = _ops(_hash_perms(64 + k * w1b), 0, 0, 0)
    assert prep + main + final == verify_ops(mode, mlen)
    return {"vf_prep": prep, "vf_main": main, "vf_final": final}


def keygen_ops(mode: int) -> int:
    k, l, _, pkb = PARAMS[mode]
    perms = 1 + 1 + 2 * (l + k) + 5 * k * l + _hash_perms(pkb)
    return _ops(perms, l, k, k * l)


def sign_essential_bytes(mode: int, mlen: int) -> int:
    sk = {2: 2528, 3: 4000, 5: 4864}[mode]
    sig = {2: 2420, 3: 3293, 5: 4595}[mode]
    return sk + mlen + sig
