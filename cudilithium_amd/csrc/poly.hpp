// Polynomial data movement for one wavefront: loads and stores in the NTT
// register layouts (A: j = 64r + l, D: j = 4l + r) and little-endian
// bit-field packing, the byte format of every packed polynomial on the path
// (t1 10-bit, t0 13-bit, eta 3/4-bit, z 18/20-bit, w1 6/4-bit; see
// keypair.cu:116-130,268-301 and sign.cu:662-683,1151-1192).
#pragma once

#include <hip/hip_runtime.h>

#include "field.hpp"
#include "ntt.hpp"
#include "params.hpp"

namespace dil {

// W-bit field j of a little-endian bit-packed array at any byte alignment.
// Reads only the bytes that hold the field.
template <int W>
__device__ __forceinline__ uint32_t field_at(const uint8_t *p, int j) {
    const int bit = j * W;
    const uint8_t *q = p + (bit >> 3);
    const int sh = bit & 7;
    constexpr int NB = (W + 7 + 7) / 8;
    uint32_t v = q[0];
#pragma unroll
    for (int b = 1; b < NB; ++b)
        if (8 * b < sh + W) v |= (uint32_t)q[b] << (8 * b);
    return (v >> sh) & ((1u << W) - 1u);
}

// Packs 8 fields of W bits (W bytes) to out, any alignment.
template <int W>
__device__ __forceinline__ void pack8_store(uint8_t *out, const uint32_t (&c)[8]) {
    constexpr int NW = (8 * W + 31) / 32;
    uint32_t acc[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) acc[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int bit = i * W, wi = bit >> 5, sh = bit & 31;
        acc[wi] |= c[i] << sh;
        if (sh + W > 32) acc[wi + 1] |= c[i] >> (32 - sh);
    }
#pragma unroll
    for (int b = 0; b < W; ++b) out[b] = (uint8_t)(acc[b >> 2] >> (8 * (b & 3)));
}

// Wave-level packing of a polynomial in layout D whose lane values are the
// W-bit fields: even lanes pack 8 consecutive coefficients into W bytes.
template <int W>
__device__ __forceinline__ void wave_pack_D(uint8_t *out, const uint32_t (&v)[4]) {
    const int l = lane_id();
    uint32_t n[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) n[r] = (uint32_t)__shfl_down((int)v[r], 1);
    if ((l & 1) == 0) {
        const uint32_t c[8] = {v[0], v[1], v[2], v[3], n[0], n[1], n[2], n[3]};
        pack8_store<W>(out + (l >> 1) * W, c);
    }
}

// Wave-level unpack of W-bit fields into layout A.
template <int W>
__device__ __forceinline__ void wave_fields_A(uint32_t (&v)[4], const uint8_t *p) {
    const int l = lane_id();
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = field_at<W>(p, 64 * r + l);
}

__device__ __forceinline__ void wave_load_A(int32_t (&v)[4], const int32_t *p) {
    const int l = lane_id();
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = p[64 * r + l];
}
__device__ __forceinline__ void wave_store_A(int32_t *p, const int32_t (&v)[4]) {
    const int l = lane_id();
#pragma unroll
    for (int r = 0; r < 4; ++r) p[64 * r + l] = v[r];
}
__device__ __forceinline__ void wave_load_D(int32_t (&v)[4], const int32_t *p) {
    const int4 t = reinterpret_cast<const int4 *>(p)[lane_id()];
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
}
__device__ __forceinline__ void wave_store_D(int32_t *p, const int32_t (&v)[4]) {
    reinterpret_cast<int4 *>(p)[lane_id()] = make_int4(v[0], v[1], v[2], v[3]);
}
// Challenge polynomial from its masks (nz[8] | neg[8], see challenge()) in
// layout A: lane l, register r holds c[64 r + l] in {-1, 0, 1}.
__device__ __forceinline__ void wave_load_challenge_A(int32_t (&v)[4], const uint32_t *masks) {
    const int l = lane_id(), sh = l & 31, hi = l >> 5;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t nzw = masks[2 * r + hi], ngw = masks[8 + 2 * r + hi];
        const int32_t nzb = (int32_t)((nzw >> sh) & 1u), ngb = (int32_t)((ngw >> sh) & 1u);
        v[r] = nzb - 2 * (nzb & ngb);
    }
}

__device__ __forceinline__ void wave_load_i8_A(int32_t (&v)[4], const int8_t *p) {
    const int l = lane_id();
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = p[64 * r + l];
}

// Wave-wide OR of a per-lane predicate.
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

}  // namespace dil
