// Batched verification (reference: gpu_verify, src/verify.cu:28-325, one
// 32-lane warp per item, A expanded into the pool and read back).
//
// Three launches:
//   vf_prep   lane per item: hint format checks + hint bitmap, tr = H(pk),
//             mu = H(tr || m), c = SampleInBall(c~)
//   vf_main   G items per workgroup: z unpack + norm check + NTT into LDS,
//             NTT(c); one lane per ExpandA entry folds accepted coefficients
//             into w[k][ctr] += mont(a, z^[l][ctr]) (LDS atomics, A never
//             stored); then per (item, k) wavefront: - c^ t1^ 2^d, INTT,
//             caddq, UseHint, pack w1'
//   vf_final  lane per item: c~' = H(mu || w1'), compare, write the verdict
// ret is 0 for a valid signature and -1 otherwise (api.cu:454-515).
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "layout.hpp"
#include "pipelines.hpp"
#include "poly.hpp"
#include "sample.hpp"

namespace dil {

template <class P>
__global__ __launch_bounds__(256) void vf_prep_kernel(VerifyArgs a) {
    using Z = Sizes<P>;
    using VL = VerifyLayout<P>;
    // serial per-lane message absorb: highest wave issue priority (see sg_mu_kernel)
    __builtin_amdgcn_s_setprio(3);
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint8_t *sig = a.sig + i * a.sig_pitch;
    uint8_t *work = a.work + i * a.work_pitch;
    int32_t *flag = reinterpret_cast<int32_t *>(work + VL::FLAG);
    if (a.key_idx && a.key_idx[i] >= a.nkeys) {
        *flag = 1;
        return;
    }
    const uint8_t *key = a.key_idx ? a.keys + a.key_idx[i] * a.key_pitch : nullptr;
    const uint8_t *pk = key ? key + KeyLayout<P>::PK : a.pk + i * a.pk_pitch;
    uint32_t *bm = reinterpret_cast<uint32_t *>(work + VL::HINT);

    // hint decoding rules (verify.cu:52-93; spec unpack_sig)
    const uint8_t *h = sig + Z::SIG_H;
#pragma unroll
    for (int w = 0; w < P::K * 8; ++w) bm[w] = 0;
    int bad = 0;
    uint32_t prev = 0;
    for (int k = 0; k < P::K && !bad; ++k) {
        const uint32_t end = h[P::OMEGA + k];
        if (end < prev || end > (uint32_t)P::OMEGA) {
            bad = 1;
            break;
        }
        for (uint32_t j = prev; j < end; ++j) {
            const uint32_t idx = h[j];
            if (j > prev && idx <= h[j - 1]) {
                bad = 1;
                break;
            }
            bm[k * 8 + (idx >> 5)] |= 1u << (idx & 31);
        }
        prev = end;
    }
    for (uint32_t j = prev; j < (uint32_t)P::OMEGA && !bad; ++j)
        if (h[j]) bad = 1;
    *flag = bad;
    if (bad) return;

    // tr = SHAKE256(pk)[0:32] (stored with a resident key); mu = SHAKE256(tr || m)[0:64]
    // (verify.cu:130-144); already in the work row when vf_mu_row_kernel ran
    if (!a.row_mu) {
        Lane64 s[25];
        uint64_t tr[4];
        if (key) {
#pragma unroll
            for (int w = 0; w < 4; ++w) tr[w] = load_le64(key + KeyLayout<P>::TR + 8 * w);
        } else {
            shake_absorb<SHAKE256_RATE>(s, [&](uint32_t w, uint32_t nb) {
                return nb == 8 ? load_le64(pk + 8 * w) : load_le_partial(pk + 8 * w, nb);
            }, (uint32_t)Z::PK_BYTES);
            squeeze_words<4>(s, tr);
        }
        const uint8_t *m = a.m.ptr(i);
        const uint32_t mlen = a.m.len(i);
        shake_absorb<SHAKE256_RATE>(s, [&](uint32_t w, uint32_t nb) -> uint64_t {
            if (w < 4) return w == 0 ? tr[0] : (w == 1 ? tr[1] : (w == 2 ? tr[2] : tr[3]));
            const uint8_t *p = m + 8 * (w - 4);
            return nb == 8 ? load_le64(p) : load_le_partial(p, nb);
        }, SEEDBYTES + mlen);
        uint64_t mu[8];
        squeeze_words<8>(s, mu);
        uint64_t *hash = reinterpret_cast<uint64_t *>(work + VL::HASH);
#pragma unroll
        for (int w = 0; w < 8; ++w) hash[w] = mu[w];
    }

    uint64_t ct[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) ct[w] = load_le64(sig + 8 * w);
    challenge<P::TAU>(ct, reinterpret_cast<uint32_t *>(work + VL::CP));
}

template <class P>
struct VfMainCfg {
    static constexpr int G = P::MODE == 2 ? 16 : (P::MODE == 3 ? 8 : 4);
};

// RES: keys are resident (A and t1^ read from the store; no accumulator);
// G items per workgroup (the compiled default VfMainCfg<P>::G, or the value
// a tiling chooses among those whose LDS fits in 160 KiB)
template <class P, bool RES, int G>
struct VfShared {
    ZetaLds zt;
    int32_t x[4][NTT_XBUF];
    int32_t zhat[G][P::L][N];
    int32_t acc[RES ? 1 : G][P::K][N];
    int32_t chat[G][N];
    int32_t fail[G];
};

template <class P, bool RES, int G>
__global__ __launch_bounds__(256) void vf_main_kernel(VerifyArgs a) {
    using Z = Sizes<P>;
    using VL = VerifyLayout<P>;
    using KL = KeyLayout<P>;
    extern __shared__ __align__(16) uint8_t smem_raw[];
    VfShared<P, RES, G> &sh = *reinterpret_cast<VfShared<P, RES, G> *>(smem_raw);
    const size_t item0 = (size_t)blockIdx.x * G;
    const int wave = threadIdx.x >> 6;
    zeta_lds_load(sh.zt);
    if constexpr (!RES)
        for (int j = threadIdx.x; j < G * P::K * N; j += blockDim.x) (&sh.acc[0][0][0])[j] = 0;
    if (threadIdx.x < G) {
        const size_t item = item0 + threadIdx.x;
        sh.fail[threadIdx.x] =
            item < a.n ? *reinterpret_cast<const int32_t *>(a.work + item * a.work_pitch + VL::FLAG) : 1;
    }
    __syncthreads();

    // z: unpack, ||z|| < gamma1 - beta, NTT (verify.cu:96-126); NTT(c)
    for (int t = wave; t < G * (P::L + 1); t += 4) {
        const int g = t / (P::L + 1), l = t % (P::L + 1);
        const size_t item = item0 + g;
        if (item >= a.n || sh.fail[g]) continue;
        int32_t v[4];
        if (l < P::L) {
            uint32_t f[4];
            wave_fields_A<Z::Z_BITS>(f, a.sig + item * a.sig_pitch + Z::SIG_Z + l * Z::POLYZ_BYTES);
            bool big = false;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = P::GAMMA1 - (int32_t)f[r];
                big |= exceeds(v[r], P::GAMMA1 - P::BETA);
            }
            if (wave_any(big)) {
                if (lane_id() == 0) sh.fail[g] = 1;
                continue;
            }
            ntt_wave(v, sh.x[wave], sh.zt);
            wave_store_D(sh.zhat[g][l], v);
        } else {
            wave_load_challenge_A(v, reinterpret_cast<const uint32_t *>(a.work + item * a.work_pitch + VL::CP));
            ntt_wave(v, sh.x[wave], sh.zt);
            wave_store_D(sh.chat[g], v);
        }
    }
    __syncthreads();

    // A z^ with A streamed from its sponges
    if constexpr (!RES)
    for (int t = threadIdx.x; t < G * P::K * P::L; t += blockDim.x) {
        const int g = t / (P::K * P::L), kl = t % (P::K * P::L);
        const int k = kl / P::L, l = kl % P::L;
        if (item0 + g >= a.n || sh.fail[g]) continue;
        const uint8_t *pk = a.pk + (item0 + g) * a.pk_pitch;
        uint64_t rho[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) rho[w] = load_le64(pk + 8 * w);
        int32_t *acc = sh.acc[g][k];
        const int32_t *zh = sh.zhat[g][l];
        uniform_poly(rho, (uint32_t)((k << 8) | l), [&](int ctr, int32_t v) { atomicAdd(&acc[ctr], mont_mul(v, zh[ctr])); });
    }
    __syncthreads();
    if (threadIdx.x < G && item0 + threadIdx.x < a.n && sh.fail[threadIdx.x])
        *reinterpret_cast<int32_t *>(a.work + (item0 + threadIdx.x) * a.work_pitch + VL::FLAG) = 1;

    // w1' = UseHint(h, caddq(INTT(A z^ - c^ t1^ 2^d))) (verify.cu:257-305)
    for (int t = wave; t < G * P::K; t += 4) {
        const int g = t / P::K, k = t % P::K;
        const size_t item = item0 + g;
        if (item >= a.n || sh.fail[g]) continue;
        uint8_t *work = a.work + item * a.work_pitch;
        int32_t t1[4], w[4], c[4];
        if constexpr (RES) {
            // A z^ from the stored A, in the order of the accumulating form (exact int32 sums)
            const uint8_t *key = a.keys + a.key_idx[item] * a.key_pitch;
            const int32_t *mat = reinterpret_cast<const int32_t *>(key + KL::MAT);
#pragma unroll
            for (int r = 0; r < 4; ++r) w[r] = 0;
            for (int l = 0; l < P::L; ++l) {
                int32_t av[4], zv[4];
                wave_load_D(av, mat + (k * P::L + l) * N);
                wave_load_D(zv, sh.zhat[g][l]);
#pragma unroll
                for (int r = 0; r < 4; ++r) w[r] += mont_mul(av[r], zv[r]);
            }
            wave_load_D(t1, reinterpret_cast<const int32_t *>(key + KL::T1H) + k * N);
        } else {
            uint32_t f[4];
            wave_fields_A<10>(f, a.pk + item * a.pk_pitch + SEEDBYTES + k * POLYT1_BYTES);
#pragma unroll
            for (int r = 0; r < 4; ++r) t1[r] = (int32_t)(f[r] << D);
            ntt_wave(t1, sh.x[wave], sh.zt);
            wave_load_D(w, sh.acc[g][k]);
        }
        wave_load_D(c, sh.chat[g]);
#pragma unroll
        for (int r = 0; r < 4; ++r) w[r] = reduce32(w[r] - mont_mul(c[r], t1[r]));
        intt_wave(w, sh.x[wave], sh.zt);
        const uint32_t *bm = reinterpret_cast<const uint32_t *>(work + VL::HINT) + k * 8;
        const int l = lane_id();
        int32_t w1[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 64 * r + l;
            const int hb = (bm[j >> 5] >> (j & 31)) & 1;
            w1[r] = use_hint<P>(caddq(w[r]), hb);
        }
        layout_A_to_D(w1, sh.x[wave]);
        uint32_t u[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) u[r] = (uint32_t)w1[r];
        wave_pack_D<Z::W1_BITS>(work + VL::HASH + CRHBYTES + k * Z::POLYW1_BYTES, u);
    }
}

template <class P>
__global__ __launch_bounds__(256) void vf_final_kernel(VerifyArgs a) {
    using Z = Sizes<P>;
    using VL = VerifyLayout<P>;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint8_t *work = a.work + i * a.work_pitch;
    int32_t verdict = -1;
    if (*reinterpret_cast<const int32_t *>(work + VL::FLAG) == 0) {
        const uint64_t *hin = reinterpret_cast<const uint64_t *>(work + VL::HASH);
        Lane64 s[25];
        shake_absorb<SHAKE256_RATE>(s, [&](uint32_t w, uint32_t) { return hin[w]; }, (uint32_t)Z::CHASH_IN_BYTES);
        uint64_t c2[4];
        squeeze_words<4>(s, c2);
        const uint8_t *sig = a.sig + i * a.sig_pitch;
        uint64_t diff = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) diff |= c2[w] ^ load_le64(sig + 8 * w);
        verdict = diff == 0 ? 0 : -1;
    }
    a.ret[i * a.ret_pitch] = verdict;
}

template <class P, int G>
constexpr bool vf_fits() {
    return sizeof(VfShared<P, false, G>) <= 160 * 1024;  // the resident form needs less
}

template <class P, bool RES, int G>
static int launch_vf_main_g(const VerifyArgs &a, hipStream_t st) {
    if constexpr (!vf_fits<P, G>()) {
        return -1;
    } else {
        const size_t smem = sizeof(VfShared<P, RES, G>);
        // the attribute is set once per kernel (a host call that need not be
        // repeated per launch; it is not a stream operation)
        static const hipError_t attr = hipFuncSetAttribute((const void *)vf_main_kernel<P, RES, G>,
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (attr != hipSuccess) return (int)attr;
        hipLaunchKernelGGL((vf_main_kernel<P, RES, G>), dim3((unsigned)((a.n + G - 1) / G)), dim3(256), smem, st, a);
        return 0;
    }
}

template <class P, bool RES>
static int launch_vf_main(const VerifyArgs &a, hipStream_t st) {
    switch (tiling_get(TILE_VERIFY_ITEMS, P::MODE, a.n)) {
        case 2: return launch_vf_main_g<P, RES, 2>(a, st);
        case 4: return launch_vf_main_g<P, RES, 4>(a, st);
        case 8: return launch_vf_main_g<P, RES, 8>(a, st);
        case 16: return launch_vf_main_g<P, RES, 16>(a, st);
        default: return launch_vf_main_g<P, RES, VfMainCfg<P>::G>(a, st);
    }
}

template <class P>
static bool verify_items_ok(int g) {
    switch (g) {
        case 0: return true;
        case 2: return vf_fits<P, 2>();
        case 4: return vf_fits<P, 4>();
        case 8: return vf_fits<P, 8>();
        case 16: return vf_fits<P, 16>();
        default: return false;
    }
}

bool verify_items_supported(int mode, int g) {
    switch (mode) {
        case 2: return verify_items_ok<Params<2>>(g);
        case 3: return verify_items_ok<Params<3>>(g);
        case 5: return verify_items_ok<Params<5>>(g);
        default: return false;
    }
}

// tr = SHAKE256(pk)[0:32] (or the resident key's) and mu = SHAKE256(tr || m)
// into the work row, two lanes per item (keccak_f1600_pair), 128 items per
// workgroup: the same bytes vf_prep writes when it hashes itself
template <class P>
__global__ __launch_bounds__(256) void vf_mu_pair_kernel(VerifyArgs a) {
    using Z = Sizes<P>;
    using VL = VerifyLayout<P>;
    __builtin_amdgcn_s_setprio(3);
    const size_t i = (size_t)blockIdx.x * 128 + (threadIdx.x >> 1);
    const uint32_t h = threadIdx.x & 1;
    if (i >= a.n) return;  // both lanes of a pair
    if (a.key_idx && a.key_idx[i] >= a.nkeys) return;  // vf_prep rejects the item
    const uint8_t *key = a.key_idx ? a.keys + a.key_idx[i] * a.key_pitch : nullptr;
    uint32_t s[25], tr[4];
    if (key) {
#pragma unroll
        for (int w = 0; w < 4; ++w) tr[w] = load_half_bytes(key + KeyLayout<P>::TR, 8 * w + 4 * h, SEEDBYTES);
    } else {
        const uint8_t *pk = a.pk + i * a.pk_pitch;
        shake_pair<SHAKE256_RATE>(s, h, (uint32_t)Z::PK_BYTES,
                                  [&](uint32_t w) { return load_half_bytes(pk, 8 * w + 4 * h, Z::PK_BYTES); });
#pragma unroll
        for (int w = 0; w < 4; ++w) tr[w] = s[w];
    }
    const uint8_t *m = a.m.ptr(i);
    const uint32_t mlen = a.m.len(i);
    shake_pair<SHAKE256_RATE>(s, h, SEEDBYTES + mlen, [&](uint32_t w) -> uint32_t {
        if (w < 4) return w == 0 ? tr[0] : (w == 1 ? tr[1] : (w == 2 ? tr[2] : tr[3]));
        return load_half_bytes(m, 8 * (w - 4) + 4 * h, mlen);
    });
    uint32_t *hash = reinterpret_cast<uint32_t *>(a.work + i * a.work_pitch + VL::HASH);
#pragma unroll
    for (int w = 0; w < 8; ++w) hash[2 * w + h] = s[w];
}

template <class P>
static int verify_impl(const VerifyArgs &a, hipStream_t st, hipEvent_t *ev) {
    if (a.n == 0) return 0;
    const unsigned b1 = (unsigned)((a.n + 255) / 256);
    int e;
    if (ev && (e = (int)hipEventRecord(ev[0], st))) return e;
    if (a.row_mu && (e = launch_verify_mu_row(P::MODE, a, st))) return e;
    // fork: mu on an auxiliary stream beside vf_prep and vf_main, joined
    // before vf_final (event fork / join; a graph capture records branches)
    AuxStreams *aux = nullptr;
    std::unique_lock<std::mutex> guard;
    VerifyArgs b = a;
    if (!a.row_mu && a.fork_mu) {
        aux = aux_streams(st);
        if (!aux) return (int)hipErrorOutOfMemory;
        guard = std::unique_lock<std::mutex>(aux->lock);
        if ((e = (int)hipEventRecord(aux->fork, st))) return e;
        if ((e = (int)hipStreamWaitEvent(aux->s[0], aux->fork, 0))) return e;
        hipLaunchKernelGGL(vf_mu_pair_kernel<P>, dim3((unsigned)((a.n + 127) / 128)), dim3(256), 0, aux->s[0], a);
        if ((e = (int)hipGetLastError())) return e;
        if ((e = (int)hipEventRecord(aux->join[0], aux->s[0]))) return e;
        b.row_mu = 1;  // vf_prep: mu is computed elsewhere
    }
    hipLaunchKernelGGL(vf_prep_kernel<P>, dim3(b1), dim3(256), 0, st, b);
    if (ev && (e = (int)hipEventRecord(ev[1], st))) return e;
    e = a.key_idx ? launch_vf_main<P, true>(a, st) : launch_vf_main<P, false>(a, st);
    if (e) return e;
    if (ev && (e = (int)hipEventRecord(ev[2], st))) return e;
    if (aux && (e = (int)hipStreamWaitEvent(st, aux->join[0], 0))) return e;
    hipLaunchKernelGGL(vf_final_kernel<P>, dim3(b1), dim3(256), 0, st, a);
    if (ev && (e = (int)hipEventRecord(ev[3], st))) return e;
    return (int)hipGetLastError();
}

int launch_verify(int mode, const VerifyArgs &a, hipStream_t st, hipEvent_t *ev) {
    switch (mode) {
        case 2: return verify_impl<Params<2>>(a, st, ev);
        case 3: return verify_impl<Params<3>>(a, st, ev);
        case 5: return verify_impl<Params<5>>(a, st, ev);
        default: return -1;
    }
}

}  // namespace dil
