"""Synthetic inputs of the benchmark configurations.

Every byte comes from xoshiro256** (Blackman-Vigna) seeded through
splitmix64 from a fixed per-configuration seed, 0xD1110000 + config number.
Structural draws (lengths, flip positions) use the Python generator below;
bulk bytes (seeds, messages) use the library's identical C generator
(dil_xoshiro_bytes) on a derived seed.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib

M64 = (1 << 64) - 1
CONFIG_SEED = {n: 0xD1110000 + n for n in range(1, 6)}


def _rotl(v, k):
    return ((v << k) | (v >> (64 - k))) & M64


class Xoshiro256ss:
    def __init__(self, seed: int):
        x = seed & M64
        self.s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            self.s.append(z ^ (z >> 31))

    def next(self) -> int:
        s = self.s
        r = (_rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return r

    def below(self, n: int) -> int:
        """Uniform in [0, n) by Lemire's multiply-shift with rejection."""
        while True:
            x = self.next()
            m = x * n
            lo = m & M64
            if lo >= n or lo >= ((1 << 64) - n) % n:
                return m >> 64

    def unit(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53

    def log_uniform(self, lo: int, hi: int) -> int:
        return int(math.floor(lo * (hi / lo) ** self.unit()))


def xoshiro_bytes(seed: int, n: int) -> np.ndarray:
    out = np.zeros(max(1, n), dtype=np.uint8)
    L = _lib.load()
    L.dil_xoshiro_bytes.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t]
    L.dil_xoshiro_bytes.restype = ctypes.c_int
    _lib.check(L.dil_xoshiro_bytes(seed & M64, out.ctypes.data, n), "xoshiro")
    return out[:n]


def config1(batch: int = 128, mlen: int = 64):
    """Dilithium2 plain keygen + sign + verify: 32-byte seeds, 64-byte messages."""
    b = xoshiro_bytes(CONFIG_SEED[1], batch * (32 + mlen))
    return b[:batch * 32].reshape(batch, 32), b[batch * 32:].reshape(batch, mlen)


def config2(npoly: int = 65536):
    """Standalone NTT: two batches of npoly x 256 int32 uniform in [-q, q)."""
    q = 8380417
    raw = xoshiro_bytes(CONFIG_SEED[2], 2 * npoly * 256 * 8).view(np.uint64)
    # uniform in [0, 2q) by multiply-shift of each 64-bit output (rejection
    # probability below 2^-40 per draw is ignored)
    v = ((raw.astype(np.float64) / 2.0 ** 64) * (2 * q)).astype(np.int64)
    v = np.minimum(v, 2 * q - 1) - q
    v = v.astype(np.int32).reshape(2, npoly, 256)
    return v[0], v[1]


def messages_loguniform(seed: int, n: int, lo: int, hi: int):
    g = Xoshiro256ss(seed)
    lens = [g.log_uniform(lo, hi) for _ in range(n)]
    flat = xoshiro_bytes(seed ^ 0x5A5A5A5A5A5A5A5A, sum(lens))
    out, pos = [], 0
    for ln in lens:
        out.append(flat[pos:pos + ln].tobytes())
        pos += ln
    return out


def config3(n: int = 16384):
    """Dilithium3 signing: keypair seeds and messages of log-uniform length in
    [32, 4096] bytes."""
    seeds = xoshiro_bytes(CONFIG_SEED[3], n * 32).reshape(n, 32)
    return seeds, messages_loguniform(CONFIG_SEED[3] + 1000, n, 32, 4096)


def config4(n: int = 65536, mlen: int = 1024, flip_fraction: float = 0.1):
    """Dilithium5 verify: seeds, 1 KiB messages and the (index, bit) of the
    signatures to corrupt (Fisher-Yates choice of round(0.1 n) items)."""
    seeds = xoshiro_bytes(CONFIG_SEED[4], n * 32).reshape(n, 32)
    msgs = xoshiro_bytes(CONFIG_SEED[4] + 1, n * mlen).reshape(n, mlen)
    g = Xoshiro256ss(CONFIG_SEED[4] + 2)
    k = int(round(flip_fraction * n))
    perm = list(range(n))
    for i in range(n - 1, 0, -1):
        j = g.below(i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    chosen = sorted(perm[:k])
    bits = [g.below(4595 * 8) for _ in chosen]
    return seeds, msgs, np.array(chosen, dtype=np.int64), np.array(bits, dtype=np.int64)


SERVE_MODES = (2, 3, 5)


def config5_keys(n_keys: int = 4096):
    """Resident keypairs of the serving mix: key k has mode SERVE_MODES[k % 3]
    and seed bytes [32 k, 32 k + 32) of the config's stream.  Returns, per
    mode, (global key ids, seeds)."""
    seeds = xoshiro_bytes(CONFIG_SEED[5], n_keys * 32).reshape(n_keys, 32)
    out = {}
    for j, m in enumerate(SERVE_MODES):
        ids = np.arange(j, n_keys, len(SERVE_MODES))
        out[m] = (ids, seeds[ids])
    return out


def config5(n_requests: int, rate_per_s: float, n_keys: int = 4096, seed_offset: int = 0,
            lo: int = 32, hi: int = 65536):
    """Serving mix: Poisson arrivals at rate_per_s (exponential gaps), ops
    keygen / sign / verify with probability 0.2 / 0.5 / 0.3, mode uniform over
    {2, 3, 5}, a resident key uniform among that mode's keys (index within
    the mode's store), messages log-uniform in [lo, hi] bytes, 32-byte seeds
    for keygen.  Returns dict of arrays / lists."""
    g = Xoshiro256ss(CONFIG_SEED[5] + 100 + seed_offset)
    per_mode = {m: len(range(j, n_keys, len(SERVE_MODES))) for j, m in enumerate(SERVE_MODES)}
    ops = np.zeros(n_requests, np.int32)
    modes = np.zeros(n_requests, np.int32)
    keys = np.zeros(n_requests, np.int64)
    lens = np.zeros(n_requests, np.int64)
    gaps = np.zeros(n_requests, np.float64)
    for i in range(n_requests):
        u = g.below(10)
        ops[i] = 0 if u < 2 else (1 if u < 7 else 2)
        modes[i] = SERVE_MODES[g.below(3)]
        keys[i] = g.below(per_mode[int(modes[i])])
        lens[i] = g.log_uniform(lo, hi) if ops[i] != 0 else 0
        gaps[i] = -math.log(1.0 - g.unit()) / rate_per_s
    flat = xoshiro_bytes(CONFIG_SEED[5] + 200 + seed_offset, int(lens.sum()))
    msgs, pos = [], 0
    for ln in lens:
        msgs.append(flat[pos:pos + int(ln)].tobytes())
        pos += int(ln)
    seeds = xoshiro_bytes(CONFIG_SEED[5] + 300 + seed_offset, n_requests * 32).reshape(n_requests, 32)
    arrival_ns = (np.cumsum(gaps) * 1e9).astype(np.uint64)
    return {"ops": ops, "modes": modes, "keys": keys, "msgs": msgs, "seeds": seeds, "arrival_ns": arrival_ns}
