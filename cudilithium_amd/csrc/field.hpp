// Exact integer arithmetic mod q = 8380417 (device side).
//
// Montgomery products follow the reference's definition
// (ntt.cuh:64-72, reduce.cuh:12): mont(a, b) = a*b*2^-32 mod q with the
// representative (a*b - t*q) / 2^32, t = lo32(a*b) * q^-1 mod 2^32.  That
// value is unique, so any instruction sequence computing it is int32
// bit-exact with the reference.
#pragma once

#include <hip/hip_runtime.h>

#include "params.hpp"

namespace dil {

// mont(a, b) for two run-time operands: 4 multiplies.
__device__ __forceinline__ int32_t mont_mul(int32_t a, int32_t b) {
    int32_t t = (int32_t)((uint32_t)a * (uint32_t)b * QINV);
    return __mulhi(a, b) - __mulhi(t, Q);
}

// mont(z, b) with a constant z whose twisted form zq = z * q^-1 mod 2^32 is
// precomputed: lo32(z*b) * q^-1 == b * zq (mod 2^32), so 3 multiplies.
__device__ __forceinline__ int32_t mont_mul_tw(int32_t b, int32_t z, int32_t zq) {
    int32_t t = (int32_t)((uint32_t)b * (uint32_t)zq);
    return __mulhi(b, z) - __mulhi(t, Q);
}

// reduce.cuh:42-48: r == a (mod q), -6283009 <= r <= 6283007.
__device__ __forceinline__ int32_t reduce32(int32_t a) {
    int32_t t = (a + (1 << 22)) >> 23;
    return a - t * Q;
}

// reduce.cuh:59-62
__device__ __forceinline__ int32_t caddq(int32_t a) { return a + ((a >> 31) & Q); }

// rounding.cuh:22-28: a = a1*2^13 + a0, -2^12 < a0 <= 2^12
__device__ __forceinline__ int32_t power2round(int32_t a, int32_t &a0) {
    int32_t a1 = (a + (1 << (D - 1)) - 1) >> D;
    a0 = a - (a1 << D);
    return a1;
}

// rounding.cuh:44-59: high/low bits of a standard representative.
template <class P>
__device__ __forceinline__ int32_t decompose(int32_t a, int32_t &a0) {
    int32_t a1 = (a + 127) >> 7;
    if constexpr (P::GAMMA2 == (Q - 1) / 32) {
        a1 = (a1 * 1025 + (1 << 21)) >> 22;
        a1 &= 15;
    } else {
        a1 = (a1 * 11275 + (1 << 23)) >> 24;
        a1 ^= ((43 - a1) >> 31) & a1;
    }
    a0 = a - a1 * 2 * P::GAMMA2;
    a0 -= (((Q - 1) / 2 - a0) >> 31) & Q;
    return a1;
}

// rounding.cuh:70-75
template <class P>
__device__ __forceinline__ int make_hint(int32_t a0, int32_t a1) {
    return (a0 > P::GAMMA2) | (a0 < -P::GAMMA2) | ((a0 == -P::GAMMA2) & (a1 != 0));
}

// rounding.cuh:86-107
template <class P>
__device__ __forceinline__ int32_t use_hint(int32_t a, int hint) {
    int32_t a0;
    int32_t a1 = decompose<P>(a, a0);
    if (hint == 0) return a1;
    if constexpr (P::GAMMA2 == (Q - 1) / 32) {
        return a0 > 0 ? (a1 + 1) & 15 : (a1 - 1) & 15;
    } else {
        if (a0 > 0) return a1 == 43 ? 0 : a1 + 1;
        return a1 == 0 ? 43 : a1 - 1;
    }
}

// poly.cuh:31-54 for one coefficient: 1 iff |a| >= bound.
__device__ __forceinline__ int exceeds(int32_t a, int32_t bound) {
    int32_t t = a >> 31;
    t = a - (t & 2 * a);
    return t >= bound;
}

}  // namespace dil
