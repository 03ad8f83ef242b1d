// Per-lane samplers over SHAKE output blocks held in registers.
//
// Each lane runs one sponge (keccak.hpp) and consumes the squeezed block
// straight from its state registers: candidate positions inside a block are
// compile-time constants (static_for), so no squeezed byte is ever written to
// memory.  Accepted coefficients go to a caller-supplied sink(ctr, value),
// which stores them (global / LDS) or folds them into a product.
//
//   uniform_poly   ExpandA entry, SHAKE128(rho || le16(nonce)), 23-bit
//                  candidates < q (sign.cu:98-163, keypair.cu:140-189)
//   eta_poly       ExpandS entry, SHAKE256(rhoprime || le16(nonce)), nibbles
//                  (keypair.cu:80-115)
//   mask_raw       ExpandMask squeeze of 5 SHAKE256 blocks (sign.cu:463-526)
//   challenge      SampleInBall (sign.cu:710-734, verify.cu:221-245)
//
// All samplers keep squeezing until the polynomial is complete, as the
// specification does; the reference stops at a fixed buffer (probability of
// a difference ~0, see the oracle's notes).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "keccak.hpp"
#include "params.hpp"

namespace dil {

template <int I, int END, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < END) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, END>(f);
    }
}

template <int H>
__device__ __forceinline__ uint32_t half_word(const Lane64 (&s)[25]) {
    if constexpr (H & 1)
        return s[H >> 1].hi;
    else
        return s[H >> 1].lo;
}

// Bits [BIT, BIT + W) of the current output block (W <= 32).
template <int BIT, int W>
__device__ __forceinline__ uint32_t block_bits(const Lane64 (&s)[25]) {
    constexpr int h = BIT >> 5, sh = BIT & 31;
    uint32_t v;
    if constexpr (sh == 0)
        v = half_word<h>(s);
    else if constexpr (sh + W <= 32)
        v = half_word<h>(s) >> sh;
    else
        v = funnel(half_word<h + 1>(s), half_word<h>(s), sh);
    if constexpr (W < 32) v &= (1u << W) - 1u;
    return v;
}

template <class Sink>
__device__ __forceinline__ void uniform_poly(const uint64_t (&rho)[4], uint32_t nonce, Sink &&sink) {
    Lane64 s[25];
    const uint64_t in[5] = {rho[0], rho[1], rho[2], rho[3], (uint64_t)(nonce & 0xFFFF)};
    shake_absorb_regs<SHAKE128_RATE, SEEDBYTES + 2>(s, in);
    int ctr = 0;
    do {
        keccak_f1600(s);
        static_for<0, SHAKE128_RATE / 3>([&](auto c) {
            constexpr int C = decltype(c)::value;
            const uint32_t t = block_bits<24 * C, 23>(s);
            if (t < (uint32_t)Q && ctr < N) {
                sink(ctr, (int32_t)t);
                ++ctr;
            }
        });
    } while (ctr < N);
}

template <int ETA, class Sink>
__device__ __forceinline__ void eta_poly(const uint64_t (&seed)[8], uint32_t nonce, Sink &&sink) {
    Lane64 s[25];
    const uint64_t in[9] = {seed[0], seed[1], seed[2], seed[3], seed[4],
                            seed[5], seed[6], seed[7], (uint64_t)(nonce & 0xFFFF)};
    shake_absorb_regs<SHAKE256_RATE, CRHBYTES + 2>(s, in);
    int ctr = 0;
    do {
        keccak_f1600(s);
        // 17 state words, 16 nibbles each, low nibble first
        static_for<0, 2 * (SHAKE256_RATE / 8)>([&](auto h) {
            uint32_t w = half_word<decltype(h)::value>(s);
#pragma unroll 2
            for (int b = 0; b < 8; ++b, w >>= 4) {
                const uint32_t t = w & 15u;
                if constexpr (ETA == 2) {
                    if (t < 15u && ctr < N) {
                        sink(ctr, 2 - (int32_t)(t - ((205u * t) >> 10) * 5u));
                        ++ctr;
                    }
                } else {
                    if (t < 9u && ctr < N) {
                        sink(ctr, 4 - (int32_t)t);
                        ++ctr;
                    }
                }
            }
        });
    } while (ctr < N);
}

// 5 squeezed SHAKE256 blocks (680 bytes) of rhoprime || le16(nonce) to out.
__device__ __forceinline__ void mask_raw(const uint64_t (&rp)[8], uint32_t nonce, uint64_t *out) {
    Lane64 s[25];
    const uint64_t in[9] = {rp[0], rp[1], rp[2], rp[3], rp[4], rp[5], rp[6], rp[7], (uint64_t)(nonce & 0xFFFF)};
    shake_absorb_regs<SHAKE256_RATE, CRHBYTES + 2>(s, in);
    for (int b = 0; b < 5; ++b) {
        keccak_f1600(s);
#pragma unroll
        for (int i = 0; i < SHAKE256_RATE / 8; ++i) out[17 * b + i] = from_lane(s[i]);
    }
}

// mask_raw with two lanes per sponge (keccak_f1600_pair): lane h absorbs
// the halves h of rhoprime's words (4-byte aligned) and writes the halves h
// of the squeezed words; the pair writes the same 680 bytes.
__device__ __forceinline__ void mask_raw_pair(const uint32_t *rhoprime, uint32_t nonce, uint32_t *out, uint32_t h) {
    uint32_t s[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) s[i] = 0u;
#pragma unroll
    for (int w = 0; w < 8; ++w) s[w] = rhoprime[2 * w + h];
    // 66 input bytes: le16(nonce) and the domain byte 0x1F in the low half of
    // word 8, the final bit in the high half of word 16
    if (h == 0) s[8] = (nonce & 0xFFFFu) | (0x1Fu << 16);
    if (h == 1) s[SHAKE256_RATE / 8 - 1] = 0x80000000u;
    for (int b = 0; b < 5; ++b) {
        keccak_f1600_pair(s, h);
#pragma unroll
        for (int i = 0; i < SHAKE256_RATE / 8; ++i) out[2 * (17 * b + i) + h] = s[i];
    }
}

// SampleInBall with the polynomial held in registers as two 256-bit masks:
// bit j of nz is set when c[j] != 0 and bit j of neg when c[j] = -1.  Each
// step (c[i] = c[b]; c[b] = +-1) becomes mask-bit moves selected without
// memory, so the byte loop has no dependent loads (the dense in-memory form
// pays a memory round trip per step).  i >= N - TAU >= 192 lies in words
// 6..7.  The masks are stored as nz[0..7] | neg[0..7] (64 bytes).
// word w of a (w < 8) by a select tree on the bits of w: no indexed access
// that the compiler could lower to scratch memory
__device__ __forceinline__ uint32_t mask_word(const uint32_t (&a)[8], uint32_t w) {
    const bool b0 = w & 1u, b1 = w & 2u, b2 = w & 4u;
    const uint32_t t0 = b0 ? a[1] : a[0], t1 = b0 ? a[3] : a[2];
    const uint32_t t2 = b0 ? a[5] : a[4], t3 = b0 ? a[7] : a[6];
    const uint32_t u0 = b1 ? t1 : t0, u1 = b1 ? t3 : t2;
    return b2 ? u1 : u0;
}

// store: false computes the masks without writing them (a lane that runs
// the same challenge as its pair partner, which writes)
template <int TAU>
__device__ __forceinline__ void challenge(const uint64_t (&ct)[4], uint32_t *masks, bool store = true) {
    static_assert(N - TAU >= 192, "challenge positions i lie in mask words 6..7");
    uint32_t nz[8], neg[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nz[k] = neg[k] = 0;
    Lane64 s[25];
    shake_absorb_regs<SHAKE256_RATE, SEEDBYTES>(s, ct);
    keccak_f1600(s);
    uint64_t signs = from_lane(s[0]);
    uint32_t i = N - TAU;
    auto step = [&](uint32_t b) {
        if (i < (uint32_t)N && b <= i) {
            const uint32_t wb = b >> 5, mb = 1u << (b & 31), mi = 1u << (i & 31);
            const bool hi7 = (i >> 5) == 7;
            // c[i] = c[b]
            const uint32_t nzb = (mask_word(nz, wb) & mb) ? mi : 0u;
            const uint32_t ngb = (mask_word(neg, wb) & mb) ? mi : 0u;
            nz[6] = hi7 ? nz[6] : ((nz[6] & ~mi) | nzb);
            nz[7] = hi7 ? ((nz[7] & ~mi) | nzb) : nz[7];
            neg[6] = hi7 ? neg[6] : ((neg[6] & ~mi) | ngb);
            neg[7] = hi7 ? ((neg[7] & ~mi) | ngb) : neg[7];
            // c[b] = (-1)^(sign bit); word wb selected by a one-hot mask
            const uint32_t sb = (uint32_t)signs & 1u ? mb : 0u, onehot = 1u << wb;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t here = 0u - ((onehot >> k) & 1u);
                nz[k] |= mb & here;
                neg[k] = (neg[k] & ~(mb & here)) | (sb & here);
            }
            signs >>= 1;
            ++i;
        }
    };
    // bytes 8.. of the first block (the first 8 are the sign bits), then
    // further blocks if ever needed; 4 bytes per 32-bit half-word
    // A lane is done once i reaches N (after ~55 of the 128 bytes on
    // average); the rest of the block is skipped as soon as every active lane
    // of the wave is done, which saves about half of the steps.
    auto bytes_of = [&](auto h) {
        if (__ballot(i < (uint32_t)N) == 0ull) return;
        uint32_t w = half_word<decltype(h)::value>(s);
#pragma unroll 1
        for (int b = 0; b < 4; ++b, w >>= 8) step(w & 0xFFu);
    };
    static_for<2, SHAKE256_RATE / 4>(bytes_of);
    while (i < (uint32_t)N) {
        keccak_f1600(s);
        static_for<0, SHAKE256_RATE / 4>(bytes_of);
    }
    if (!store) return;
    uint4 *o = reinterpret_cast<uint4 *>(masks);
    o[0] = make_uint4(nz[0], nz[1], nz[2], nz[3]);
    o[1] = make_uint4(nz[4], nz[5], nz[6], nz[7]);
    o[2] = make_uint4(neg[0], neg[1], neg[2], neg[3]);
    o[3] = make_uint4(neg[4], neg[5], neg[6], neg[7]);
}

// challenge with the squeeze in the pair layout (keccak_f1600_pair): lane h
// of the pair holds halves h of the block, the partner's half comes by one
// DPP move per 4 bytes consumed; both lanes sample the same polynomial and
// only lane h = 0 stores it.  Both lanes of every pair must be active.
template <int TAU>
__device__ __forceinline__ void challenge_pair(const uint64_t (&ct)[4], uint32_t *masks, uint32_t hh) {
    static_assert(N - TAU >= 192, "challenge positions i lie in mask words 6..7");
    uint32_t nz[8], neg[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nz[k] = neg[k] = 0;
    uint32_t s[25];
#pragma unroll
    for (int w = 0; w < 25; ++w) s[w] = 0u;
#pragma unroll
    for (int w = 0; w < 4; ++w) s[w] = hh ? (uint32_t)(ct[w] >> 32) : (uint32_t)ct[w];
    if (hh == 0) s[4] = 0x1Fu;  // 32 input bytes: domain byte at byte 32
    if (hh == 1) s[SHAKE256_RATE / 8 - 1] = 0x80000000u;
    keccak_f1600_pair(s, hh);
    uint64_t signs = pair_word(s, 0, hh);
    uint32_t i = N - TAU;
    auto step = [&](uint32_t b) {
        if (i < (uint32_t)N && b <= i) {
            const uint32_t wb = b >> 5, mb = 1u << (b & 31), mi = 1u << (i & 31);
            const bool hi7 = (i >> 5) == 7;
            // c[i] = c[b]
            const uint32_t nzb = (mask_word(nz, wb) & mb) ? mi : 0u;
            const uint32_t ngb = (mask_word(neg, wb) & mb) ? mi : 0u;
            nz[6] = hi7 ? nz[6] : ((nz[6] & ~mi) | nzb);
            nz[7] = hi7 ? ((nz[7] & ~mi) | nzb) : nz[7];
            neg[6] = hi7 ? neg[6] : ((neg[6] & ~mi) | ngb);
            neg[7] = hi7 ? ((neg[7] & ~mi) | ngb) : neg[7];
            // c[b] = (-1)^(sign bit); word wb selected by a one-hot mask
            const uint32_t sb = (uint32_t)signs & 1u ? mb : 0u, onehot = 1u << wb;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t here = 0u - ((onehot >> k) & 1u);
                nz[k] |= mb & here;
                neg[k] = (neg[k] & ~(mb & here)) | (sb & here);
            }
            signs >>= 1;
            ++i;
        }
    };
    // bytes 8.. of the first block (the first 8 are the sign bits), then
    // further blocks if ever needed; 4 bytes per 32-bit half-word
    // A lane is done once i reaches N (after ~55 of the 128 bytes on
    // average); the rest of the block is skipped as soon as every active lane
    // of the wave is done, which saves about half of the steps.
    auto bytes_of = [&](auto h) {
        if (__ballot(i < (uint32_t)N) == 0ull) return;  // wave-uniform: the DPP below runs in every lane
        constexpr int J = decltype(h)::value;
        const uint32_t own = s[J / 2], other = pair_other(own);
        uint32_t w = (uint32_t)(J & 1) == hh ? own : other;
#pragma unroll 1
        for (int b = 0; b < 4; ++b, w >>= 8) step(w & 0xFFu);
    };
    static_for<2, SHAKE256_RATE / 4>(bytes_of);
    while (i < (uint32_t)N) {  // same trip count in both lanes of a pair
        keccak_f1600_pair(s, hh);
        static_for<0, SHAKE256_RATE / 4>(bytes_of);
    }
    if (hh != 0) return;
    uint4 *o = reinterpret_cast<uint4 *>(masks);
    o[0] = make_uint4(nz[0], nz[1], nz[2], nz[3]);
    o[1] = make_uint4(nz[4], nz[5], nz[6], nz[7]);
    o[2] = make_uint4(neg[0], neg[1], neg[2], neg[3]);
    o[3] = make_uint4(neg[4], neg[5], neg[6], neg[7]);
}

// Squeezes the first output block after absorption and returns words [0, NW).
template <int NW>
__device__ __forceinline__ void squeeze_words(Lane64 (&s)[25], uint64_t (&out)[NW]) {
    keccak_f1600(s);
#pragma unroll
    for (int i = 0; i < NW; ++i) out[i] = from_lane(s[i]);
}

}  // namespace dil
