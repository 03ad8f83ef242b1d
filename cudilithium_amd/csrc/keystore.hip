// Expansion of resident keys (KeyLayout rows).
//
// The reference recomputes ExpandA on every signature (sign.cu:98-163) and
// every verification (verify.cu:146-200), and unpacks + transforms s1, s2,
// t0 on every signature (sign.cu:306-403).  For keys that stay resident
// these depend only on the key, so they are computed once here:
//   ks_expand  lane per A entry   A[k][l] = RejUniform(SHAKE128(rho||l||k))
//   ks_prep    wave per poly      s1^, s2^, t0^ (with the secret key) and
//                                 t1^ = NTT(t1 << d) from the public key
//   ks_tr      lane per key       tr = SHAKE256(pk)[0:32], flags
// rho is read from the public key (its first 32 bytes equal the secret
// key's), so verify-only keys expand without a secret key.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "sign_stages.hpp"

namespace dil {

constexpr int KS_WAVES = 4;

template <class P>
SignArgs key_args(uint8_t *keys, size_t pitch, size_t n, size_t sk_off) {
    SignArgs a{};
    a.sk = keys + sk_off;
    a.sk_pitch = pitch;
    a.rows = keys;
    a.row_pitch = pitch;
    a.off_ints = KeyLayout<P>::MAT;
    a.n = n;
    return a;
}

template <class P>
__global__ __launch_bounds__(256) void ks_expand_kernel(SignArgs a) {
    constexpr int KL = P::K * P::L;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * KL) return;
    const int e = (int)(t % KL);
    stage_expand_a<P>(a, t / KL, e / P::L, e % P::L);
}

// polys p < L + 2K: s1 | s2 | t0 (secret key); p >= L + 2K: t1 row k
template <class P>
__global__ __launch_bounds__(64 * KS_WAVES) void ks_prep_kernel(SignArgs a, int with_sk) {
    using KY = KeyLayout<P>;
    constexpr int NS = P::L + 2 * P::K, NP = NS + P::K;
    __shared__ ZetaLds zt;
    __shared__ int32_t x[KS_WAVES][NTT_XBUF];
    zeta_lds_load(zt);
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const size_t total = a.n * NP, stride = (size_t)gridDim.x * KS_WAVES;
    for (size_t t = (size_t)blockIdx.x * KS_WAVES + w; t < total; t += stride) {
        const size_t i = t / NP;
        const int p = (int)(t % NP);
        if (p < NS) {
            if (with_sk) {
                stage_prep<P>(a, i, p, x[w], zt);
            } else {
                // a verify-only key holds no secret: its s1^ / s2^ / t0^ rows are
                // cleared so that nothing of a row's previous key survives
                const int32_t zero[4] = {0, 0, 0, 0};
                wave_store_D(reinterpret_cast<int32_t *>(a.rows + i * a.row_pitch + KY::S1H) + p * N, zero);
            }
            continue;
        }
        const int k = p - NS;
        uint8_t *key = a.rows + i * a.row_pitch;
        uint32_t f[4];
        wave_fields_A<10>(f, key + KY::PK + SEEDBYTES + k * POLYT1_BYTES);
        int32_t v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = (int32_t)(f[r] << D);
        ntt_wave(v, x[w], zt);
        wave_store_D(reinterpret_cast<int32_t *>(key + KY::T1H) + k * N, v);
    }
}

template <class P>
__global__ __launch_bounds__(256) void ks_tr_kernel(uint8_t *keys, size_t pitch, size_t n, int with_sk) {
    using Z = Sizes<P>;
    using KY = KeyLayout<P>;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t *key = keys + i * pitch;
    const uint8_t *pk = key + KY::PK;
    Lane64 s[25];
    shake_absorb<SHAKE256_RATE>(s, [&](uint32_t w, uint32_t nb) {
        return nb == 8 ? load_le64(pk + 8 * w) : load_le_partial(pk + 8 * w, nb);
    }, (uint32_t)Z::PK_BYTES);
    uint64_t tr[4];
    squeeze_words<4>(s, tr);
#pragma unroll
    for (int w = 0; w < 4; ++w) reinterpret_cast<uint64_t *>(key + KY::TR)[w] = tr[w];  // 256-B aligned
    *reinterpret_cast<uint32_t *>(key + KY::FLAGS) = with_sk ? KEY_HAS_SECRET : 0u;
    if (!with_sk) {
        // verify-only key: clear the secret-key bytes a previous key left in the row
        uint32_t *skw = reinterpret_cast<uint32_t *>(key + KY::SK);  // 256-B aligned
        for (int w = 0; w < (Z::SK_BYTES + 3) / 4; ++w) skw[w] = 0;
    }
}

template <class P>
static int key_expand_impl(uint8_t *keys, size_t pitch, size_t n, int with_sk, hipStream_t st) {
    if (n == 0) return 0;
    constexpr int KL = P::K * P::L, NP = P::L + 3 * P::K;
    const SignArgs a = key_args<P>(keys, pitch, n, with_sk ? KeyLayout<P>::SK : KeyLayout<P>::PK);
    hipLaunchKernelGGL(ks_tr_kernel<P>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, pitch, n, with_sk);
    hipLaunchKernelGGL(ks_expand_kernel<P>, dim3((unsigned)((n * KL + 255) / 256)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(ks_prep_kernel<P>, dim3(default_grid_blocks(n * NP, KS_WAVES)), dim3(64 * KS_WAVES), 0, st,
                       a, with_sk);
    return (int)hipGetLastError();
}

int launch_key_expand(int mode, uint8_t *keys, size_t key_pitch, size_t n, int with_sk, hipStream_t st) {
    switch (mode) {
        case 2: return key_expand_impl<Params<2>>(keys, key_pitch, n, with_sk, st);
        case 3: return key_expand_impl<Params<3>>(keys, key_pitch, n, with_sk, st);
        case 5: return key_expand_impl<Params<5>>(keys, key_pitch, n, with_sk, st);
        default: return -1;
    }
}

}  // namespace dil
