// Forward / inverse NTT mod q, one polynomial per 64-lane wavefront.
//
// The butterfly DAG is the reference's (src/ntt.cu:58-207, the same DAG as
// the pq-crystals C ntt / invntt_tomont): forward Cooley-Tukey levels
// len = 128..1 with zetas[1..255], inverse Gentleman-Sande levels len = 1..128
// with -zetas[255..1] and a final Montgomery scaling by 41978.  No
// intermediate reductions, so the int32 outputs are bit-identical to the
// reference for identical int32 inputs.
//
// Wave64 layout: lane l (0..63) holds 4 coefficients v[0..3].  Two levels are
// done in registers at a time (radix 4), with three LDS transposes between
// the four register layouts
//     A: j = 64 r + l             (levels len 128, 64)
//     B: j = 64 (l>>4) + 16 r + (l&15)   (len 32, 16)
//     C: j = 16 (l>>2) + 4 r + (l&3)     (len 8, 4)
//     D: j = 4 l + r              (len 2, 1)
// Forward takes layout A and returns D (4 consecutive coefficients per lane,
// one 16-byte store); inverse takes D and returns A.  All zetas come from an
// LDS table with their twisted form zq = z * q^-1 mod 2^32, so each
// butterfly costs 3 multiplies.
#pragma once

#include <hip/hip_runtime.h>

#include "field.hpp"
#include "params.hpp"

namespace dil {

constexpr int32_t MONT_2_DIV_N = 41978;  // 2^64 / 256 mod q (ntt.cu:203)

constexpr int brv8(int i) {
    int r = 0;
    for (int b = 0; b < 8; ++b) r |= ((i >> b) & 1) << (7 - b);
    return r;
}

struct ZetaConsts {
    int32_t z[N];
    int32_t zq[N];
};

// zetas[i] = centred(2^32 * 1753^brv8(i) mod q), zetas[0] = 0 (ntt.cu:24-56)
constexpr ZetaConsts make_zetas() {
    ZetaConsts t{};
    const uint64_t mont = (1ull << 32) % (uint64_t)Q;
    for (int i = 1; i < N; ++i) {
        uint64_t p = 1, b = 1753;
        for (int e = brv8(i); e; e >>= 1) {
            if (e & 1) p = p * b % Q;
            b = b * b % Q;
        }
        int64_t v = (int64_t)(p * mont % Q);
        if (v > (Q - 1) / 2) v -= Q;
        t.z[i] = (int32_t)v;
        t.zq[i] = (int32_t)((uint32_t)t.z[i] * QINV);
    }
    return t;
}

constexpr ZetaConsts kZeta = make_zetas();
static_assert(kZeta.z[1] == 25847 && kZeta.z[2] == -2608894 && kZeta.z[255] == 1976782,
              "zetas closed form");

// LDS copy of the zetas used by every NTT in a workgroup: {z, zq} pairs so
// that one 64-bit LDS read fetches both
struct ZetaLds {
    int2 zz[N];
};

__device__ inline const ZetaConsts &zeta_consts() {
    static __constant__ ZetaConsts tab = kZeta;
    return tab;
}

__device__ __forceinline__ void zeta_lds_load(ZetaLds &t) {
    const ZetaConsts &c = zeta_consts();
    for (int i = threadIdx.x; i < N; i += blockDim.x) t.zz[i] = make_int2(c.z[i], c.zq[i]);
}

// Orders LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Exchange buffer per wavefront: 256 coefficients, padded so that each
// transpose is free of LDS bank conflicts.
constexpr int NTT_XBUF = N + N / 16;
__device__ __forceinline__ int xpad(int j) { return j + (j >> 4); }

__device__ __forceinline__ void ct_bfly(int32_t &a, int32_t &b, const ZetaLds &zt, int k) {
    const int2 z = zt.zz[k];
    const int32_t t = mont_mul_tw(b, z.x, z.y);
    b = a - t;
    a = a + t;
}

__device__ __forceinline__ void gs_bfly(int32_t &a, int32_t &b, const ZetaLds &zt, int k) {
    const int2 z = zt.zz[k];
    const int32_t t = a;
    a = t + b;
    b = mont_mul_tw(t - b, -z.x, (int32_t)(0u - (uint32_t)z.y));
}

// A <-> B without LDS: swaps register bit 1 with lane bit 5 (upper half of
// one register with the lower half of its partner) and register bit 0 with
// lane bit 4 (odd 16-lane rows with even rows), gfx950's v_permlane32_swap
// and v_permlane16_swap.
__device__ __forceinline__ void swap_AB(int32_t (&v)[4]) {
    auto p0 = __builtin_amdgcn_permlane32_swap(v[0], v[2], false, false);
    auto p1 = __builtin_amdgcn_permlane32_swap(v[1], v[3], false, false);
    v[0] = p0[0];
    v[2] = p0[1];
    v[1] = p1[0];
    v[3] = p1[1];
    auto q0 = __builtin_amdgcn_permlane16_swap(v[0], v[1], false, false);
    auto q1 = __builtin_amdgcn_permlane16_swap(v[2], v[3], false, false);
    v[0] = q0[0];
    v[1] = q0[1];
    v[2] = q1[0];
    v[3] = q1[1];
}

__device__ __forceinline__ int idx_A(int l, int r) { return 64 * r + l; }
__device__ __forceinline__ int idx_B(int l, int r) { return 64 * (l >> 4) + 16 * r + (l & 15); }
__device__ __forceinline__ int idx_C(int l, int r) { return 16 * (l >> 2) + 4 * r + (l & 3); }
__device__ __forceinline__ int idx_D(int l, int r) { return 4 * l + r; }

template <class From, class To>
__device__ __forceinline__ void transpose(int32_t (&v)[4], int32_t *x, int l, From from, To to) {
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) x[xpad(from(l, r))] = v[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = x[xpad(to(l, r))];
}

// Forward NTT: v in layout A on entry, layout D on exit.
__device__ __forceinline__ void ntt_wave(int32_t (&v)[4], int32_t *x, const ZetaLds &zt) {
    const int l = lane_id();
    // len 128 (register bit 1), len 64 (register bit 0)
    ct_bfly(v[0], v[2], zt, 1);
    ct_bfly(v[1], v[3], zt, 1);
    ct_bfly(v[0], v[1], zt, 2);
    ct_bfly(v[2], v[3], zt, 3);
    swap_AB(v);
    {
        const int h = l >> 4;  // j >> 6
        ct_bfly(v[0], v[2], zt, 4 + h);
        ct_bfly(v[1], v[3], zt, 4 + h);
        ct_bfly(v[0], v[1], zt, 8 + 2 * h);
        ct_bfly(v[2], v[3], zt, 9 + 2 * h);
    }
    transpose(v, x, l, idx_B, idx_C);
    {
        const int h = l >> 2;  // j >> 4
        ct_bfly(v[0], v[2], zt, 16 + h);
        ct_bfly(v[1], v[3], zt, 16 + h);
        ct_bfly(v[0], v[1], zt, 32 + 2 * h);
        ct_bfly(v[2], v[3], zt, 33 + 2 * h);
    }
    transpose(v, x, l, idx_C, idx_D);
    ct_bfly(v[0], v[2], zt, 64 + l);
    ct_bfly(v[1], v[3], zt, 64 + l);
    ct_bfly(v[0], v[1], zt, 128 + 2 * l);
    ct_bfly(v[2], v[3], zt, 129 + 2 * l);
}

// Inverse NTT with the Montgomery factor (invntt_tomont): layout D on entry,
// layout A on exit.  Level len = 2^m uses -zetas[2^(8-m) - 1 - (j >> (m+1))].
__device__ __forceinline__ void intt_wave(int32_t (&v)[4], int32_t *x, const ZetaLds &zt) {
    const int l = lane_id();
    gs_bfly(v[0], v[1], zt, 255 - 2 * l);
    gs_bfly(v[2], v[3], zt, 254 - 2 * l);
    gs_bfly(v[0], v[2], zt, 127 - l);
    gs_bfly(v[1], v[3], zt, 127 - l);
    transpose(v, x, l, idx_D, idx_C);
    {
        const int h = l >> 2;
        gs_bfly(v[0], v[1], zt, 63 - 2 * h);
        gs_bfly(v[2], v[3], zt, 62 - 2 * h);
        gs_bfly(v[0], v[2], zt, 31 - h);
        gs_bfly(v[1], v[3], zt, 31 - h);
    }
    transpose(v, x, l, idx_C, idx_B);
    {
        const int h = l >> 4;
        gs_bfly(v[0], v[1], zt, 15 - 2 * h);
        gs_bfly(v[2], v[3], zt, 14 - 2 * h);
        gs_bfly(v[0], v[2], zt, 7 - h);
        gs_bfly(v[1], v[3], zt, 7 - h);
    }
    swap_AB(v);
    gs_bfly(v[0], v[1], zt, 3);
    gs_bfly(v[2], v[3], zt, 2);
    gs_bfly(v[0], v[2], zt, 1);
    gs_bfly(v[1], v[3], zt, 1);
    constexpr int32_t fq = (int32_t)((uint32_t)MONT_2_DIV_N * QINV);
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = mont_mul_tw(v[r], MONT_2_DIV_N, fq);
}

// Converts between layouts through the exchange buffer (no arithmetic).
__device__ __forceinline__ void layout_D_to_A(int32_t (&v)[4], int32_t *x) {
    transpose(v, x, lane_id(), idx_D, idx_A);
}
__device__ __forceinline__ void layout_A_to_D(int32_t (&v)[4], int32_t *x) {
    transpose(v, x, lane_id(), idx_A, idx_D);
}

}  // namespace dil
