// Message hashing with one sponge per 16-lane row (keccak_row.hpp), for
// batches of long messages, where the serial absorb of the longest message
// (floor((32 + mlen) / 136) + 1 permutations) sets the batch's latency.
//
//   sg_mu_row   mu = SHAKE256(tr || m) and rhoprime = SHAKE256(key || mu)
//               (or the caller's rhoprime), the fixed-stage part of signing
//               that stage_mu computes with one lane per item
//               (api.cu:137-153)
//   vf_mu_row   tr = SHAKE256(pk) (or the resident key's) and
//               mu = SHAKE256(tr || m) into the verify work row, before
//               vf_prep (verify.cu:130-144), which then skips them
//
// Four items per wavefront, sixteen per 256-thread workgroup.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "keccak_row.hpp"
#include "layout.hpp"
#include "pipelines.hpp"
#include "sign_stages.hpp"

namespace dil {

constexpr int ROW_ITEMS_PER_BLOCK = 16;

// a[x] holds half r.h of output word x + 5 y: lanes with word < nwords
// store their half to dst (32-bit stores; dst 4-byte aligned)
__device__ __forceinline__ void row_store_words(const uint32_t (&a)[5], const RowLane &r, uint32_t *dst,
                                                uint32_t nwords) {
#pragma unroll
    for (int x = 0; x < 5; ++x) {
        const uint32_t w = row_word(r, x);
        if (r.active && w < nwords) dst[2 * w + r.h] = a[x];
    }
}

// shares output words [0, nwords) of the group through LDS (the state's pi
// area, free after the permutation): returns half h of word w
struct RowShare {
    uint32_t base;
    __device__ void put(uint32_t *lds, const uint32_t (&a)[5], const RowLane &r, uint32_t nwords) const {
#pragma unroll
        for (int x = 0; x < 5; ++x) {
            const uint32_t w = row_word(r, x);
            if (r.active && w < nwords) lds[base + 2 * w + r.h] = a[x];
        }
    }
    __device__ uint32_t get(const uint32_t *lds, uint32_t w, uint32_t h) const { return lds[base + 2 * w + h]; }
};

template <class P>
__global__ __launch_bounds__(256) void sg_mu_row_kernel(SignArgs a) {
    using Z = Sizes<P>;
    using SLy = SignLayout<P>;
    __shared__ uint32_t lds[4 * 4 * ROW_LDS_WORDS];
    __builtin_amdgcn_s_setprio(3);  // a serial chain: issue priority (see sg_mu_kernel)
    const uint32_t wave = threadIdx.x >> 6;
    const RowLane r = row_lane_setup(lds, wave * 4 * ROW_LDS_WORDS);
    const size_t i = (size_t)blockIdx.x * ROW_ITEMS_PER_BLOCK + wave * 4 + ((uint32_t)lane_id() >> 4);
    if (i >= a.n) return;  // a whole 16-lane row: no DPP partner of another item is affected
    const uint8_t *sk = a.sk + sk_index(a, i) * a.sk_pitch;
    const uint8_t *m = a.m.ptr(i);
    const uint32_t mlen = a.m.len(i);
    const uint32_t h = r.h;
    uint32_t s[5];
    // mu = SHAKE256(tr || m)
    shake_row<SHAKE256_RATE>(s, r, lds, SEEDBYTES + mlen, [&](uint32_t w) -> uint32_t {
        if (w < 4) return load_half_bytes(sk + Z::SK_TR, 8 * w + 4 * h, SEEDBYTES);
        return load_half_bytes(m, 8 * (w - 4) + 4 * h, mlen);
    });
    uint8_t *row = a.rows + i * a.row_pitch;
    row_store_words(s, r, reinterpret_cast<uint32_t *>(row + SLy::MU), 8);
    if (a.rnd) {
        const uint8_t *rnd = a.rnd + i * a.rnd_pitch;
        const uint32_t g = (uint32_t)lane_id() & 15;  // 16 lanes x 4 bytes
        reinterpret_cast<uint32_t *>(row + SLy::RHOPRIME)[g] = load_half_bytes(rnd, 4 * g, CRHBYTES);
        return;
    }
    // rhoprime = SHAKE256(key || mu): mu's words cross rows, through LDS
    const RowShare sh{wave * 4 * ROW_LDS_WORDS + ((uint32_t)lane_id() >> 4) * ROW_LDS_WORDS};
    sh.put(lds, s, r, 8);
    wave_lds_sync();
    uint32_t t[5];
    shake_row<SHAKE256_RATE>(t, r, lds, SEEDBYTES + CRHBYTES, [&](uint32_t w) -> uint32_t {
        if (w < 4) return load_half_bytes(sk + Z::SK_KEY, 8 * w + 4 * h, SEEDBYTES);
        return sh.get(lds, w - 4, h);
    });
    row_store_words(t, r, reinterpret_cast<uint32_t *>(row + SLy::RHOPRIME), 8);
}

template <class P>
__global__ __launch_bounds__(256) void vf_mu_row_kernel(VerifyArgs a) {
    using Z = Sizes<P>;
    using VL = VerifyLayout<P>;
    __shared__ uint32_t lds[4 * 4 * ROW_LDS_WORDS];
    __builtin_amdgcn_s_setprio(3);
    const uint32_t wave = threadIdx.x >> 6;
    const RowLane r = row_lane_setup(lds, wave * 4 * ROW_LDS_WORDS);
    const size_t i = (size_t)blockIdx.x * ROW_ITEMS_PER_BLOCK + wave * 4 + ((uint32_t)lane_id() >> 4);
    if (i >= a.n) return;
    if (a.key_idx && a.key_idx[i] >= a.nkeys) return;  // vf_prep rejects the item
    const uint32_t h = r.h;
    const RowShare sh{wave * 4 * ROW_LDS_WORDS + ((uint32_t)lane_id() >> 4) * ROW_LDS_WORDS};
    uint32_t s[5];
    const uint8_t *key = a.key_idx ? a.keys + a.key_idx[i] * a.key_pitch : nullptr;
    if (!key) {
        // tr = SHAKE256(pk)[0:32]
        const uint8_t *pk = a.pk + i * a.pk_pitch;
        shake_row<SHAKE256_RATE>(s, r, lds, (uint32_t)Z::PK_BYTES,
                                 [&](uint32_t w) { return load_half_bytes(pk, 8 * w + 4 * h, Z::PK_BYTES); });
        sh.put(lds, s, r, 4);
        wave_lds_sync();
    }
    const uint8_t *m = a.m.ptr(i);
    const uint32_t mlen = a.m.len(i);
    shake_row<SHAKE256_RATE>(s, r, lds, SEEDBYTES + mlen, [&](uint32_t w) -> uint32_t {
        if (w < 4) return key ? load_half_bytes(key + KeyLayout<P>::TR, 8 * w + 4 * h, SEEDBYTES) : sh.get(lds, w, h);
        return load_half_bytes(m, 8 * (w - 4) + 4 * h, mlen);
    });
    row_store_words(s, r, reinterpret_cast<uint32_t *>(a.work + i * a.work_pitch + VL::HASH), 8);
}

template <class P>
static int sign_mu_row_impl(const SignArgs &a, hipStream_t st) {
    if (a.n == 0) return 0;
    hipLaunchKernelGGL(sg_mu_row_kernel<P>, dim3((unsigned)((a.n + ROW_ITEMS_PER_BLOCK - 1) / ROW_ITEMS_PER_BLOCK)),
                       dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

template <class P>
static int verify_mu_row_impl(const VerifyArgs &a, hipStream_t st) {
    if (a.n == 0) return 0;
    hipLaunchKernelGGL(vf_mu_row_kernel<P>, dim3((unsigned)((a.n + ROW_ITEMS_PER_BLOCK - 1) / ROW_ITEMS_PER_BLOCK)),
                       dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

int launch_sign_mu_row(int mode, const SignArgs &a, hipStream_t st) {
    switch (mode) {
        case 2: return sign_mu_row_impl<Params<2>>(a, st);
        case 3: return sign_mu_row_impl<Params<3>>(a, st);
        case 5: return sign_mu_row_impl<Params<5>>(a, st);
        default: return -1;
    }
}

int launch_verify_mu_row(int mode, const VerifyArgs &a, hipStream_t st) {
    switch (mode) {
        case 2: return verify_mu_row_impl<Params<2>>(a, st);
        case 3: return verify_mu_row_impl<Params<3>>(a, st);
        case 5: return verify_mu_row_impl<Params<5>>(a, st);
        default: return -1;
    }
}

}  // namespace dil
