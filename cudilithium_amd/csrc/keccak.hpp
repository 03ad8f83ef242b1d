// Keccak-f[1600] and SHAKE128/256 with one sponge state per lane.
//
// The reference spreads one state over 25 lanes of a 32-wide warp and moves
// data with shuffles (fips202.cuh:104-147).  On CDNA4 each of the 64 lanes
// of a wavefront instead owns a whole state in 50 VGPRs: no cross-lane
// traffic, all 64 lanes busy, and the round maps onto gfx950's 3-input
// v_bitop3_b32 (theta parities, theta update, chi) and v_alignbit_b32 (64-bit
// rotations as two funnel shifts).  Batched work supplies the independent
// sponges: ExpandA alone has K*L of them per key.
//
// The round constants and rotation offsets are generated at compile time from
// the FIPS 202 definitions (LFSR rc(t); offsets (t+1)(t+2)/2 along the
// (x,y) -> (y, 2x+3y) walk), the same values the reference tabulates in
// src/fips202/fips202.cu:13-34.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dil {

struct KeccakConsts {
    uint64_t rc[24];
    int rho[25];  // rotation of lane x + 5y
    int pi[25];   // destination lane of lane x + 5y under pi: (x,y) -> (y, 2x+3y)
};

constexpr int keccak_rc_bit(int t) {
    int r = 1;
    for (int i = 1; i <= t % 255; ++i) {
        r <<= 1;
        if (r & 0x100) r ^= 0x171;
    }
    return r & 1;
}

constexpr KeccakConsts make_keccak_consts() {
    KeccakConsts k{};
    for (int ir = 0; ir < 24; ++ir) {
        uint64_t rc = 0;
        for (int j = 0; j < 7; ++j)
            if (keccak_rc_bit(j + 7 * ir)) rc |= 1ull << ((1 << j) - 1);
        k.rc[ir] = rc;
    }
    int x = 1, y = 0;
    k.rho[0] = 0;
    for (int t = 0; t < 24; ++t) {
        k.rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx;
        y = ny;
    }
    for (int xx = 0; xx < 5; ++xx)
        for (int yy = 0; yy < 5; ++yy) k.pi[xx + 5 * yy] = yy + 5 * ((2 * xx + 3 * yy) % 5);
    return k;
}

constexpr KeccakConsts kKeccak = make_keccak_consts();
static_assert(kKeccak.rc[0] == 1 && kKeccak.rc[23] == 0x8000000080008008ull, "FIPS 202 iota");
static_assert(kKeccak.rho[1] == 1 && kKeccak.rho[24] == 14, "FIPS 202 rho");

__constant__ uint64_t kKeccakRC[24] = {
    kKeccak.rc[0],  kKeccak.rc[1],  kKeccak.rc[2],  kKeccak.rc[3],  kKeccak.rc[4],  kKeccak.rc[5],
    kKeccak.rc[6],  kKeccak.rc[7],  kKeccak.rc[8],  kKeccak.rc[9],  kKeccak.rc[10], kKeccak.rc[11],
    kKeccak.rc[12], kKeccak.rc[13], kKeccak.rc[14], kKeccak.rc[15], kKeccak.rc[16], kKeccak.rc[17],
    kKeccak.rc[18], kKeccak.rc[19], kKeccak.rc[20], kKeccak.rc[21], kKeccak.rc[22], kKeccak.rc[23]};

// bitop3 truth tables: result bit = imm[(a << 2) | (b << 1) | c]
constexpr uint32_t bitop3_imm(int f000, int f001, int f010, int f011, int f100, int f101, int f110,
                              int f111) {
    return (uint32_t)(f000 | f001 << 1 | f010 << 2 | f011 << 3 | f100 << 4 | f101 << 5 | f110 << 6 |
                      f111 << 7);
}
constexpr uint32_t kXor3 = bitop3_imm(0, 1, 1, 0, 1, 0, 0, 1);  // a ^ b ^ c
constexpr uint32_t kChi = bitop3_imm(0, 1, 0, 0, 1, 0, 1, 1);   // a ^ (~b & c)

struct Lane64 {
    uint32_t lo, hi;
};

template <uint32_t IMM>
__device__ __forceinline__ uint32_t bitop3(uint32_t a, uint32_t b, uint32_t c) {
    return (uint32_t)__builtin_amdgcn_bitop3_b32(a, b, c, IMM);
}

// funnel shift: low 32 bits of ({hi, lo} >> s), s in [1, 31]
__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
    return (uint32_t)__builtin_amdgcn_alignbit(hi, lo, s);
}

__device__ __forceinline__ Lane64 xor3(Lane64 a, Lane64 b, Lane64 c) {
    return {bitop3<kXor3>(a.lo, b.lo, c.lo), bitop3<kXor3>(a.hi, b.hi, c.hi)};
}

__device__ __forceinline__ Lane64 chi(Lane64 a, Lane64 b, Lane64 c) {
    return {bitop3<kChi>(a.lo, b.lo, c.lo), bitop3<kChi>(a.hi, b.hi, c.hi)};
}

// rotate left by a compile-time constant n in [0, 64)
template <int n>
__device__ __forceinline__ Lane64 rotl(Lane64 x) {
    if constexpr (n == 0) {
        return x;
    } else if constexpr (n == 32) {
        return {x.hi, x.lo};
    } else if constexpr (n < 32) {
        return {funnel(x.lo, x.hi, 32 - n), funnel(x.hi, x.lo, 32 - n)};
    } else {
        return {funnel(x.hi, x.lo, 64 - n), funnel(x.lo, x.hi, 64 - n)};
    }
}

template <int I>
__device__ __forceinline__ void rho_pi_step(const Lane64 (&a)[25], Lane64 (&b)[25]) {
    b[kKeccak.pi[I]] = rotl<kKeccak.rho[I]>(a[I]);
    if constexpr (I + 1 < 25) rho_pi_step<I + 1>(a, b);
}

__device__ __forceinline__ void keccak_round(Lane64 (&a)[25], uint64_t rc) {
    Lane64 c[5], b[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = xor3(xor3(a[x], a[x + 5], a[x + 10]), a[x + 15], a[x + 20]);
#pragma unroll
    for (int x = 0; x < 5; ++x) {
        Lane64 r = rotl<1>(c[(x + 1) % 5]);
#pragma unroll
        for (int y = 0; y < 5; ++y) a[x + 5 * y] = xor3(a[x + 5 * y], c[(x + 4) % 5], r);
    }
    rho_pi_step<0>(a, b);
#pragma unroll
    for (int y = 0; y < 5; ++y)
#pragma unroll
        for (int x = 0; x < 5; ++x)
            a[x + 5 * y] = chi(b[x + 5 * y], b[(x + 1) % 5 + 5 * y], b[(x + 2) % 5 + 5 * y]);
    a[0].lo ^= (uint32_t)rc;
    a[0].hi ^= (uint32_t)(rc >> 32);
}

__device__ __forceinline__ void keccak_f1600(Lane64 (&a)[25]) {
#pragma unroll 2
    for (int r = 0; r < 24; ++r) keccak_round(a, kKeccakRC[r]);
}

__device__ __forceinline__ Lane64 to_lane(uint64_t v) { return {(uint32_t)v, (uint32_t)(v >> 32)}; }
__device__ __forceinline__ uint64_t from_lane(Lane64 v) { return ((uint64_t)v.hi << 32) | v.lo; }

__device__ __forceinline__ void xor_word(Lane64 &a, uint64_t v) {
    a.lo ^= (uint32_t)v;
    a.hi ^= (uint32_t)(v >> 32);
}

// Loads nbytes (1..8) little-endian from p; bytes past nbytes are zero.
__device__ __forceinline__ uint64_t load_le_partial(const uint8_t *p, uint32_t nbytes) {
    uint64_t w = 0;
    for (uint32_t i = 0; i < nbytes; ++i) w |= (uint64_t)p[i] << (8 * i);
    return w;
}

// Loads 8 bytes little-endian from p of any alignment.
__device__ __forceinline__ uint64_t load_le64(const uint8_t *p) {
    if (((uintptr_t)p & 7) == 0) return *reinterpret_cast<const uint64_t *>(p);
    if (((uintptr_t)p & 3) == 0) {
        const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
        return (uint64_t)q[0] | ((uint64_t)q[1] << 32);
    }
    return load_le_partial(p, 8);
}

// SHAKE absorb of a byte stream delivered by a word source: get(w, n) returns
// word w of the input (n <= 8 valid bytes; the rest zero).  Leaves the state
// padded (domain byte 0x1F, final bit 0x80) and ready for squeezing, i.e. the
// next keccak_f1600 produces the first output block (fips202.cuh:176-214).
template <int RATE, class Get>
__device__ __forceinline__ void shake_absorb(Lane64 (&s)[25], Get get, uint32_t len) {
    constexpr int RW = RATE / 8;
#pragma unroll
    for (int i = 0; i < 25; ++i) s[i] = {0u, 0u};
    uint32_t base = 0;
    // full blocks, the next block's words loaded before the current block's
    // permutation so that their memory latency overlaps it
    uint64_t cur[RW];
    if (len >= (uint32_t)RATE) {
#pragma unroll
        for (int i = 0; i < RW; ++i) cur[i] = get(i, 8u);
    }
    while (len >= (uint32_t)RATE) {
#pragma unroll
        for (int i = 0; i < RW; ++i) xor_word(s[i], cur[i]);
        base += RW;
        len -= RATE;
        if (len >= (uint32_t)RATE) {
#pragma unroll
            for (int i = 0; i < RW; ++i) cur[i] = get(base + i, 8u);
        }
        keccak_f1600(s);
    }
    const uint32_t nw = len >> 3, rem = len & 7;
#pragma unroll
    for (int i = 0; i < RW; ++i) {
        uint64_t w = 0;
        if ((uint32_t)i < nw)
            w = get(base + i, 8u);
        else if ((uint32_t)i == nw)
            w = (rem ? get(base + i, rem) : 0ull) | (0x1Full << (8 * rem));
        xor_word(s[i], w);
    }
    s[RW - 1].hi ^= 0x80000000u;
}

// Absorb of an input held in registers: `words` (little-endian) covering
// exactly LEN < RATE bytes with zero bytes past LEN.
template <int RATE, int LEN, int NW>
__device__ __forceinline__ void shake_absorb_regs(Lane64 (&s)[25], const uint64_t (&words)[NW]) {
    constexpr int RW = RATE / 8;
    static_assert(NW <= RW && LEN < RATE && (LEN + 7) / 8 <= NW, "single block only");
#pragma unroll
    for (int i = 0; i < 25; ++i) s[i] = {0u, 0u};
#pragma unroll
    for (int i = 0; i < NW; ++i) xor_word(s[i], words[i]);
    xor_word(s[LEN >> 3], 0x1Full << (8 * (LEN & 7)));
    s[RW - 1].hi ^= 0x80000000u;
}

// ---------------------------------------------------------------- lane pairs
// Keccak-f[1600] over a pair of lanes (2j, 2j + 1): lane 2j + h holds half h
// (h = 0: bits 0-31, h = 1: bits 32-63) of all 25 state words in 25 VGPRs.
// Bitwise steps (theta parities and update, chi, iota) need only the own
// half; a 64-bit rotation needs the other half, read from the partner lane
// with one DPP move (quad_perm [1,0,3,2]), and is then one funnel shift, the
// same instruction for both halves.  A round is about 125 VALU instructions
// against 180 for one state per lane, and a wave holds 32 states instead of
// 64: for a handful of serial sponges (the challenge hash of a few dozen
// signing attempts) the chain is a third shorter, with the same lane count.

// value of the partner lane (lane ^ 1)
__device__ __forceinline__ uint32_t pair_other(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm(1, 0, 3, 2)
}

// own half of rotl64(word, R) for a word whose own / other halves are given
template <int R>
__device__ __forceinline__ uint32_t rotl_half(uint32_t own, uint32_t other) {
    if constexpr (R == 0) {
        return own;
    } else if constexpr (R == 32) {
        return other;
    } else if constexpr (R < 32) {
        return funnel(own, other, 32 - R);
    } else {
        return funnel(other, own, 64 - R);
    }
}

template <int I>
__device__ __forceinline__ void rho_pi_pair_step(const uint32_t (&a)[25], uint32_t (&b)[25]) {
    if constexpr (kKeccak.rho[I] == 0)
        b[kKeccak.pi[I]] = a[I];
    else
        b[kKeccak.pi[I]] = rotl_half<kKeccak.rho[I]>(a[I], pair_other(a[I]));
    if constexpr (I + 1 < 25) rho_pi_pair_step<I + 1>(a, b);
}

__device__ __forceinline__ void keccak_round_pair(uint32_t (&a)[25], uint32_t rc_half) {
    uint32_t c[5], co[5], b[25];
#pragma unroll
    for (int x = 0; x < 5; ++x)
        c[x] = bitop3<kXor3>(bitop3<kXor3>(a[x], a[x + 5], a[x + 10]), a[x + 15], a[x + 20]);
#pragma unroll
    for (int x = 0; x < 5; ++x) co[x] = pair_other(c[x]);
#pragma unroll
    for (int x = 0; x < 5; ++x) {
        const uint32_t r = rotl_half<1>(c[(x + 1) % 5], co[(x + 1) % 5]);
#pragma unroll
        for (int y = 0; y < 5; ++y) a[x + 5 * y] = bitop3<kXor3>(a[x + 5 * y], c[(x + 4) % 5], r);
    }
    rho_pi_pair_step<0>(a, b);
#pragma unroll
    for (int y = 0; y < 5; ++y)
#pragma unroll
        for (int x = 0; x < 5; ++x)
            a[x + 5 * y] = bitop3<kChi>(b[x + 5 * y], b[(x + 1) % 5 + 5 * y], b[(x + 2) % 5 + 5 * y]);
    a[0] ^= rc_half;
}

// h: this lane's half (lane & 1); both lanes of the pair must be active.
// Rounds unrolled DIL_PAIR_UNROLL at a time (default 2: the signing loop's
// code size matters more than the round-constant loads).
#ifndef DIL_PAIR_UNROLL
#define DIL_PAIR_UNROLL 2
#endif
__device__ __forceinline__ void keccak_f1600_pair(uint32_t (&a)[25], uint32_t h) {
#pragma unroll DIL_PAIR_UNROLL
    for (int r = 0; r < 24; ++r) {
        const uint64_t rc = kKeccakRC[r];
        keccak_round_pair(a, h ? (uint32_t)(rc >> 32) : (uint32_t)rc);
    }
}

// SHAKE absorb of LEN bytes (a multiple of 8, known at compile time) in the
// pair layout: get(w) returns this lane's half of input word w.  Leaves the
// state padded for the first squeeze (the pair counterpart of shake_absorb).
template <int RATE, int LEN, class Get>
__device__ __forceinline__ void shake_absorb_pair(uint32_t (&s)[25], Get get, uint32_t h) {
    constexpr int RW = RATE / 8, NWORDS = LEN / 8, FULL = NWORDS / RW, REST = NWORDS % RW;
    static_assert(LEN % 8 == 0, "whole words");
#pragma unroll
    for (int i = 0; i < 25; ++i) s[i] = 0u;
    uint32_t cur[RW];
    if constexpr (FULL > 0) {
#pragma unroll
        for (int i = 0; i < RW; ++i) cur[i] = get(i);
    }
#pragma unroll 1
    for (int blk = 0; blk < FULL; ++blk) {
#pragma unroll
        for (int i = 0; i < RW; ++i) s[i] ^= cur[i];
        if (blk + 1 < FULL) {
#pragma unroll
            for (int i = 0; i < RW; ++i) cur[i] = get((uint32_t)((blk + 1) * RW + i));
        }
        keccak_f1600_pair(s, h);
    }
#pragma unroll
    for (int i = 0; i < REST; ++i) s[i] ^= get((uint32_t)(FULL * RW + i));
    // domain byte 0x1F at byte 0 of word REST (low half), final bit 63 of word RW - 1
    if (h == 0) s[REST] ^= 0x1Fu;
    if (h == 1) s[RW - 1] ^= 0x80000000u;
}

// 4 bytes little-endian at p + off, zero past len (any alignment)
__device__ __forceinline__ uint32_t load_half_bytes(const uint8_t *p, uint32_t off, uint32_t len) {
    if (off + 4 <= len && ((reinterpret_cast<uintptr_t>(p) + off) & 3) == 0)
        return *reinterpret_cast<const uint32_t *>(p + off);
    uint32_t v = 0;
    for (uint32_t b = 0; b < 4 && off + b < len; ++b) v |= (uint32_t)p[off + b] << (8 * b);
    return v;
}

// SHAKE absorb of a byte stream of run-time length `len` in the pair layout,
// plus the first squeeze permutation: get_half(w) returns this lane's half of
// input word w (bytes past the end zero; only called for w < ceil(len / 8)).
// Afterwards s[w] holds half h of output word w of the first block.  Both
// lanes of the pair run the same trip count (same item).
template <int RATE, class GetHalf>
__device__ __forceinline__ void shake_pair(uint32_t (&s)[25], uint32_t h, uint32_t len, GetHalf get_half) {
    constexpr uint32_t RW = RATE / 8;
#pragma unroll
    for (int i = 0; i < 25; ++i) s[i] = 0u;
    const uint32_t nwords = (len + 7) / 8, nfull = len / RATE;
    auto fetch = [&](uint32_t blk, uint32_t i) -> uint32_t {
        const uint32_t w = blk * RW + i;
        return w < nwords ? get_half(w) : 0u;
    };
    uint32_t cur[RW];
#pragma unroll
    for (uint32_t i = 0; i < RW; ++i) cur[i] = fetch(0, i);
#pragma unroll 1
    for (uint32_t blk = 0; blk < nfull; ++blk) {
#pragma unroll
        for (uint32_t i = 0; i < RW; ++i) s[i] ^= cur[i];
#pragma unroll
        for (uint32_t i = 0; i < RW; ++i) cur[i] = fetch(blk + 1, i);  // overlaps the permutation
        keccak_f1600_pair(s, h);
    }
    // last (partial) block: domain byte 0x1F at byte rem, final bit 0x80 at byte RATE - 1
    const uint32_t rem = len - nfull * RATE;
#pragma unroll
    for (uint32_t i = 0; i < RW; ++i) {
        uint32_t v = cur[i];
        if (i == rem / 8 && h == (rem % 8) / 4) v ^= 0x1Fu << (8 * (rem % 4));
        if (i == RW - 1 && h == 1) v ^= 0x80000000u;
        s[i] ^= v;
    }
    keccak_f1600_pair(s, h);
}

// full 64-bit word i of the state (both lanes of the pair get it)
__device__ __forceinline__ uint64_t pair_word(const uint32_t (&s)[25], int i, uint32_t h) {
    const uint32_t own = s[i], other = pair_other(own);
    const uint32_t lo = h ? other : own, hi = h ? own : other;
    return ((uint64_t)hi << 32) | lo;
}

// Byte i of the current output block (i compile-time after unrolling).
__device__ __forceinline__ uint32_t state_byte(const Lane64 (&s)[25], int i) {
    const Lane64 w = s[i >> 3];
    const int b = i & 7;
    return (b < 4 ? (w.lo >> (8 * b)) : (w.hi >> (8 * (b - 4)))) & 0xFF;
}

}  // namespace dil
