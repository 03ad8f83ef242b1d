// Dilithium round-3 v3.1 parameter sets as compile-time traits.
//
// One library serves all three modes: kernels are templated on a Params<MODE>
// type instead of being compiled once per -DDILITHIUM_MODE as the reference
// does (include/params.h:14-73, src/CMakeLists.txt:3-16).
#pragma once

#include <cstddef>
#include <cstdint>

namespace dil {

constexpr int N = 256;
constexpr int32_t Q = 8380417;
constexpr uint32_t QINV = 58728449u;  // q^-1 mod 2^32 (reduce.cuh:12)
constexpr int D = 13;
constexpr int SEEDBYTES = 32;
constexpr int CRHBYTES = 64;
constexpr int SHAKE128_RATE = 168;
constexpr int SHAKE256_RATE = 136;
constexpr int POLYT1_BYTES = 320;
constexpr int POLYT0_BYTES = 416;

constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }

template <int MODE>
struct Params;

template <>
struct Params<2> {
    static constexpr int MODE = 2, K = 4, L = 4, ETA = 2, TAU = 39, BETA = 78, OMEGA = 80;
    static constexpr int32_t GAMMA1 = 1 << 17, GAMMA2 = (Q - 1) / 88;
};
template <>
struct Params<3> {
    static constexpr int MODE = 3, K = 6, L = 5, ETA = 4, TAU = 49, BETA = 196, OMEGA = 55;
    static constexpr int32_t GAMMA1 = 1 << 19, GAMMA2 = (Q - 1) / 32;
};
template <>
struct Params<5> {
    static constexpr int MODE = 5, K = 8, L = 7, ETA = 2, TAU = 60, BETA = 120, OMEGA = 75;
    static constexpr int32_t GAMMA1 = 1 << 19, GAMMA2 = (Q - 1) / 32;
};

// Sizes derived from a parameter set (params.h:49-73).
template <class P>
struct Sizes {
    static constexpr int ETA_BITS = P::ETA == 2 ? 3 : 4;
    static constexpr int Z_BITS = P::GAMMA1 == (1 << 17) ? 18 : 20;
    static constexpr int W1_BITS = P::GAMMA2 == (Q - 1) / 88 ? 6 : 4;
    static constexpr int POLYETA_BYTES = N * ETA_BITS / 8;
    static constexpr int POLYZ_BYTES = N * Z_BITS / 8;
    static constexpr int POLYW1_BYTES = N * W1_BITS / 8;
    static constexpr int POLYVECH_BYTES = P::OMEGA + P::K;
    static constexpr int PK_BYTES = SEEDBYTES + P::K * POLYT1_BYTES;
    static constexpr int SK_BYTES = 3 * SEEDBYTES + (P::L + P::K) * POLYETA_BYTES + P::K * POLYT0_BYTES;
    static constexpr int SIG_BYTES = SEEDBYTES + P::L * POLYZ_BYTES + POLYVECH_BYTES;
    // offsets inside sk: rho | key | tr | s1 | s2 | t0
    static constexpr int SK_KEY = SEEDBYTES;
    static constexpr int SK_TR = 2 * SEEDBYTES;
    static constexpr int SK_S1 = 3 * SEEDBYTES;
    static constexpr int SK_S2 = SK_S1 + P::L * POLYETA_BYTES;
    static constexpr int SK_T0 = SK_S2 + P::K * POLYETA_BYTES;
    // offsets inside sig: c~ | z | h
    static constexpr int SIG_Z = SEEDBYTES;
    static constexpr int SIG_H = SEEDBYTES + P::L * POLYZ_BYTES;
    // c~ hash input: mu | w1 packed
    static constexpr int W1PK_BYTES = P::K * POLYW1_BYTES;
    static constexpr int CHASH_IN_BYTES = CRHBYTES + W1PK_BYTES;
    // ExpandMask squeezes 5 SHAKE256 blocks per polynomial (sign.cu:28)
    static constexpr int YRAW_BYTES = 5 * SHAKE256_RATE;
};

}  // namespace dil
