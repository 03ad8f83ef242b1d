// Batched key generation (reference: gpu_keypair, src/keypair.cu:35-311, one
// 32-lane warp per key with ExpandA streamed through shared memory).
//
// Four launches, each with the lane mapping its work wants:
//   kg_seed   lane per key     seed (explicit or SHAKE128(le64(ctr))), then
//                              rho | rhoprime | key = SHAKE256(seed)
//   kg_eta    lane per s-poly  ExpandS into int8 scratch, pack eta into sk
//   kg_main   G keys per workgroup: NTT(s1) into LDS, then one lane per
//             ExpandA entry (G*K*L sponges) folds each accepted coefficient
//             straight into t^[k][ctr] += mont(a, s1^[l][ctr]) with LDS
//             atomics -- A is never stored; the sum is exact int32 so the
//             accumulation order does not matter.  Then INTT, + s2, caddq,
//             Power2Round and packing of t1 / t0, one wavefront per row.
//   kg_tr     lane per key     tr = SHAKE256(pk)[0:32]
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "layout.hpp"
#include "pipelines.hpp"
#include "poly.hpp"
#include "sample.hpp"

namespace dil {

__device__ __forceinline__ void store_le64(uint8_t *p, uint64_t v) {
    if (((uintptr_t)p & 7) == 0) {
        *reinterpret_cast<uint64_t *>(p) = v;
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i));
    }
}

template <class P>
__global__ __launch_bounds__(256) void kg_seed_kernel(KeygenArgs a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    uint64_t seed[4];
    Lane64 s[25];
    if (a.seeds) {
        const uint8_t *sp = a.seeds + i * a.seed_pitch;
#pragma unroll
        for (int w = 0; w < 4; ++w) seed[w] = load_le64(sp + 8 * w);
    } else {
        // keypair.cu:50-60: seed = SHAKE128(le64(rand_index + 2i + 1))[0:32]
        const uint64_t in[1] = {(uint64_t)a.rand_index + 2 * (uint64_t)i + 1};
        shake_absorb_regs<SHAKE128_RATE, 8>(s, in);
        squeeze_words<4>(s, seed);
    }
    uint64_t out[16];
    shake_absorb_regs<SHAKE256_RATE, SEEDBYTES>(s, seed);
    squeeze_words<16>(s, out);
    uint8_t *pk = a.pk + i * a.pk_pitch;
    uint8_t *sk = a.sk + i * a.sk_pitch;
    uint8_t *rp = a.scratch + i * a.scratch_pitch + KeypairLayout<P>::RHOPRIME - KeypairLayout<P>::SCRATCH;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        store_le64(pk + 8 * w, out[w]);
        store_le64(sk + 8 * w, out[w]);
        store_le64(sk + SEEDBYTES + 8 * w, out[12 + w]);
    }
#pragma unroll
    for (int w = 0; w < 8; ++w) store_le64(rp + 8 * w, out[4 + w]);
}

template <class P>
__global__ __launch_bounds__(256) void kg_eta_kernel(KeygenArgs a) {
    using Z = Sizes<P>;
    constexpr int NP = P::L + P::K;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * NP) return;
    const size_t i = t / NP;
    const int p = (int)(t % NP);
    uint8_t *scr = a.scratch + i * a.scratch_pitch;
    const uint8_t *rp = scr + KeypairLayout<P>::RHOPRIME - KeypairLayout<P>::SCRATCH;
    int8_t *s8 = reinterpret_cast<int8_t *>(scr + KeypairLayout<P>::S8 - KeypairLayout<P>::SCRATCH) + p * N;
    uint64_t seed[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) seed[w] = *reinterpret_cast<const uint64_t *>(rp + 8 * w);
    // The coefficients arrive in order: they are collected into 16-byte
    // stores of the int8 copy (kg_main reads it) and into
    // the packed eta - s bit stream of the secret key (keypair.cu:116-130),
    // written a 32-bit word at a time (bytes when the row is not 4-aligned).
    // Lanes write different rows, so one store per coefficient made every
    // store instruction touch 64 cache lines for a single byte each.
    uint8_t *dst = a.sk + i * a.sk_pitch + Z::SK_S1 + p * Z::POLYETA_BYTES;
    const bool dst4 = (reinterpret_cast<uintptr_t>(dst) & 3) == 0;
    const bool s8_16 = (reinterpret_cast<uintptr_t>(s8) & 15) == 0;  // else 8-aligned (work pitch % 8 == 0)
    uint32_t cur = 0, w0 = 0, w1 = 0, w2 = 0;  // int8 bytes: the current word and the last three
    uint64_t bits = 0;
    int nbits = 0, nout = 0;
    // nonce l for s1[l], L + k for s2[k] (keypair.cu:84-89, 202-207)
    eta_poly<P::ETA>(seed, (uint32_t)p, [&](int ctr, int32_t v) {
        cur = (cur >> 8) | ((uint32_t)(uint8_t)(int8_t)v << 24);
        if ((ctr & 3) == 3) {
            if ((ctr & 15) == 15) {
                if (s8_16) {
                    reinterpret_cast<uint4 *>(s8)[ctr >> 4] = make_uint4(w0, w1, w2, cur);
                } else {
                    reinterpret_cast<uint2 *>(s8)[2 * (ctr >> 4)] = make_uint2(w0, w1);
                    reinterpret_cast<uint2 *>(s8)[2 * (ctr >> 4) + 1] = make_uint2(w2, cur);
                }
            }
            w0 = w1;
            w1 = w2;
            w2 = cur;
        }
        bits |= (uint64_t)(uint32_t)(P::ETA - v) << nbits;
        nbits += Z::ETA_BITS;
        if (nbits >= 32) {
            const uint32_t word = (uint32_t)bits;
            if (dst4) {
                reinterpret_cast<uint32_t *>(dst)[nout] = word;
            } else {
#pragma unroll
                for (int b = 0; b < 4; ++b) dst[4 * nout + b] = (uint8_t)(word >> (8 * b));
            }
            ++nout;
            bits >>= 32;
            nbits -= 32;
        }
    });
}

// keys per workgroup of kg_main: the compiled default, and the values a
// tiling may choose (those whose LDS fits in 160 KiB)
template <class P>
struct KgMainCfg {
    static constexpr int G = P::MODE == 2 ? 16 : (P::MODE == 3 ? 8 : 4);
};

template <class P, int G>
struct KgShared {
    ZetaLds zt;
    int32_t x[4][NTT_XBUF];
    int32_t shat[G][P::L][N];
    int32_t acc[G][P::K][N];
};

template <class P, int G>
__global__ __launch_bounds__(256) void kg_main_kernel(KeygenArgs a) {
    using Z = Sizes<P>;
    using KL = KeypairLayout<P>;
    extern __shared__ __align__(16) uint8_t smem_raw[];
    KgShared<P, G> &sh = *reinterpret_cast<KgShared<P, G> *>(smem_raw);
    const size_t item0 = (size_t)blockIdx.x * G;
    const int wave = threadIdx.x >> 6;
    zeta_lds_load(sh.zt);
    for (int j = threadIdx.x; j < G * P::K * N; j += blockDim.x) (&sh.acc[0][0][0])[j] = 0;
    __syncthreads();

    // NTT(s1) for the G keys
    for (int t = wave; t < G * P::L; t += 4) {
        const int g = t / P::L, l = t % P::L;
        if (item0 + g >= a.n) continue;
        const int8_t *s8 = reinterpret_cast<const int8_t *>(a.scratch + (item0 + g) * a.scratch_pitch + KL::S8 -
                                                            KL::SCRATCH);
        int32_t v[4];
        wave_load_i8_A(v, s8 + l * N);
        ntt_wave(v, sh.x[wave], sh.zt);
        wave_store_D(sh.shat[g][l], v);
    }
    __syncthreads();

    // ExpandA: one sponge per (key, k, l), accepted coefficients folded into t^
    for (int t = threadIdx.x; t < G * P::K * P::L; t += blockDim.x) {
        const int g = t / (P::K * P::L), kl = t % (P::K * P::L);
        const int k = kl / P::L, l = kl % P::L;
        if (item0 + g >= a.n) continue;
        const uint8_t *pk = a.pk + (item0 + g) * a.pk_pitch;
        uint64_t rho[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) rho[w] = load_le64(pk + 8 * w);
        int32_t *acc = sh.acc[g][k];
        const int32_t *s1 = sh.shat[g][l];
        uniform_poly(rho, (uint32_t)((k << 8) | l), [&](int ctr, int32_t v) { atomicAdd(&acc[ctr], mont_mul(v, s1[ctr])); });
    }
    __syncthreads();

    // t = INTT(reduce32(t^)) + s2, caddq, Power2Round, pack (keypair.cu:251-301)
    for (int t = wave; t < G * P::K; t += 4) {
        const int g = t / P::K, k = t % P::K;
        const size_t item = item0 + g;
        if (item >= a.n) continue;
        int32_t v[4];
        wave_load_D(v, sh.acc[g][k]);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = reduce32(v[r]);
        intt_wave(v, sh.x[wave], sh.zt);
        const int8_t *s2 = reinterpret_cast<const int8_t *>(a.scratch + item * a.scratch_pitch + KL::S8 - KL::SCRATCH) +
                           (P::L + k) * N;
        int32_t s2v[4], t1[4], t0[4];
        wave_load_i8_A(s2v, s2);
#pragma unroll
        for (int r = 0; r < 4; ++r) t1[r] = power2round(caddq(v[r] + s2v[r]), t0[r]);
        layout_A_to_D(t1, sh.x[wave]);
        layout_A_to_D(t0, sh.x[wave]);
        uint32_t f1[4], f0[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            f1[r] = (uint32_t)t1[r];
            f0[r] = (uint32_t)((1 << (D - 1)) - t0[r]);
        }
        wave_pack_D<10>(a.pk + item * a.pk_pitch + SEEDBYTES + k * POLYT1_BYTES, f1);
        wave_pack_D<13>(a.sk + item * a.sk_pitch + Z::SK_T0 + k * POLYT0_BYTES, f0);
    }
}

template <class P>
__global__ __launch_bounds__(256) void kg_tr_kernel(KeygenArgs a) {
    using Z = Sizes<P>;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint8_t *pk = a.pk + i * a.pk_pitch;
    Lane64 s[25];
    shake_absorb<SHAKE256_RATE>(s, [&](uint32_t w, uint32_t nb) {
        return nb == 8 ? load_le64(pk + 8 * w) : load_le_partial(pk + 8 * w, nb);
    }, (uint32_t)Z::PK_BYTES);
    uint64_t tr[4];
    squeeze_words<4>(s, tr);
    uint8_t *sk = a.sk + i * a.sk_pitch + Z::SK_TR;
#pragma unroll
    for (int w = 0; w < 4; ++w) store_le64(sk + 8 * w, tr[w]);
}

// tr = SHAKE256(pk)[0:32] with two lanes per key (keccak_f1600_pair): the
// serial absorb of the public key (10 / 15 / 20 permutations) is 2/3 as long
template <class P>
__global__ __launch_bounds__(256) void kg_tr_pair_kernel(KeygenArgs a) {
    using Z = Sizes<P>;
    const size_t i = (size_t)blockIdx.x * 128 + (threadIdx.x >> 1);
    const uint32_t h = threadIdx.x & 1;
    if (i >= a.n) return;  // both lanes of a pair
    const uint8_t *pk = a.pk + i * a.pk_pitch;
    uint32_t s[25];
    shake_pair<SHAKE256_RATE>(s, h, (uint32_t)Z::PK_BYTES,
                              [&](uint32_t w) { return load_half_bytes(pk, 8 * w + 4 * h, Z::PK_BYTES); });
    uint8_t *sk = a.sk + i * a.sk_pitch + Z::SK_TR;
    const bool al = (reinterpret_cast<uintptr_t>(sk) & 3) == 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint8_t *q = sk + 8 * w + 4 * h;
        if (al) {
            *reinterpret_cast<uint32_t *>(q) = s[w];
        } else {
#pragma unroll
            for (int b = 0; b < 4; ++b) q[b] = (uint8_t)(s[w] >> (8 * b));
        }
    }
}

template <class P, int G>
constexpr bool kg_fits() {
    return sizeof(KgShared<P, G>) <= 160 * 1024;
}

template <class P, int G>
static int launch_kg_main(const KeygenArgs &a, hipStream_t st) {
    if constexpr (!kg_fits<P, G>()) {
        return -1;
    } else {
        const size_t smem = sizeof(KgShared<P, G>);
        // set once per kernel: a host call, not a stream operation
        static const hipError_t attr = hipFuncSetAttribute((const void *)kg_main_kernel<P, G>,
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (attr != hipSuccess) return (int)attr;
        hipLaunchKernelGGL((kg_main_kernel<P, G>), dim3((unsigned)((a.n + G - 1) / G)), dim3(256), smem, st, a);
        return 0;
    }
}

template <class P>
static int keygen_impl(const KeygenArgs &a, hipStream_t st) {
    if (a.n == 0) return 0;
    constexpr int NP = P::L + P::K;
    const unsigned b1 = (unsigned)((a.n + 255) / 256);
    hipLaunchKernelGGL(kg_seed_kernel<P>, dim3(b1), dim3(256), 0, st, a);
    hipLaunchKernelGGL(kg_eta_kernel<P>, dim3((unsigned)((a.n * NP + 255) / 256)), dim3(256), 0, st, a);
    int e;
    switch (tiling_get(TILE_KEYGEN_ITEMS, P::MODE, a.n)) {
        case 2: e = launch_kg_main<P, 2>(a, st); break;
        case 4: e = launch_kg_main<P, 4>(a, st); break;
        case 8: e = launch_kg_main<P, 8>(a, st); break;
        case 16: e = launch_kg_main<P, 16>(a, st); break;
        default: e = launch_kg_main<P, KgMainCfg<P>::G>(a, st); break;
    }
    if (e) return e;
#ifndef DIL_KG_TR_PAIR
#define DIL_KG_TR_PAIR 1  // 16,384 keys: 1-1.5 % faster keygen than one lane per key
#endif
    if (DIL_KG_TR_PAIR)
        hipLaunchKernelGGL(kg_tr_pair_kernel<P>, dim3((unsigned)((a.n + 127) / 128)), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(kg_tr_kernel<P>, dim3(b1), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

template <class P>
static bool keygen_items_ok(int g) {
    switch (g) {
        case 0: return true;
        case 2: return kg_fits<P, 2>();
        case 4: return kg_fits<P, 4>();
        case 8: return kg_fits<P, 8>();
        case 16: return kg_fits<P, 16>();
        default: return false;
    }
}

bool keygen_items_supported(int mode, int g) {
    switch (mode) {
        case 2: return keygen_items_ok<Params<2>>(g);
        case 3: return keygen_items_ok<Params<3>>(g);
        case 5: return keygen_items_ok<Params<5>>(g);
        default: return false;
    }
}

int launch_keygen(int mode, const KeygenArgs &a, hipStream_t st) {
    switch (mode) {
        case 2: return keygen_impl<Params<2>>(a, st);
        case 3: return keygen_impl<Params<3>>(a, st);
        case 5: return keygen_impl<Params<5>>(a, st);
        default: return -1;
    }
}

}  // namespace dil
