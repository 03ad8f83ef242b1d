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

// SampleInBall into a dense int8 polynomial in memory owned by this lane.
template <int TAU>
__device__ __forceinline__ void challenge(const uint64_t (&ct)[4], int8_t *cp) {
    int4 *c4 = reinterpret_cast<int4 *>(cp);
#pragma unroll
    for (int i = 0; i < N / 16; ++i) c4[i] = make_int4(0, 0, 0, 0);
    Lane64 s[25];
    shake_absorb_regs<SHAKE256_RATE, SEEDBYTES>(s, ct);
    keccak_f1600(s);
    uint64_t signs = from_lane(s[0]);
    int i = N - TAU;
    auto step = [&](uint32_t b) {
        if (i < N && b <= (uint32_t)i) {
            cp[i] = cp[b];
            cp[b] = (signs & 1) ? (int8_t)-1 : (int8_t)1;
            signs >>= 1;
            ++i;
        }
    };
    // bytes 8.. of the first block (the first 8 are the sign bits), then
    // further blocks if ever needed; 4 bytes per 32-bit half-word
    auto bytes_of = [&](auto h) {
        uint32_t w = half_word<decltype(h)::value>(s);
#pragma unroll 1
        for (int b = 0; b < 4; ++b, w >>= 8) step(w & 0xFFu);
    };
    static_for<2, SHAKE256_RATE / 4>(bytes_of);
    while (i < N) {
        keccak_f1600(s);
        static_for<0, SHAKE256_RATE / 4>(bytes_of);
    }
}

// Squeezes the first output block after absorption and returns words [0, NW).
template <int NW>
__device__ __forceinline__ void squeeze_words(Lane64 (&s)[25], uint64_t (&out)[NW]) {
    keccak_f1600(s);
#pragma unroll
    for (int i = 0; i < NW; ++i) out[i] = from_lane(s[i]);
}

}  // namespace dil
