// Batched NTT kernels: forward, inverse, and the fused
// forward -> pointwise Montgomery -> inverse pipeline.
//
// Reference counterparts: ntt_radix8_kernel / intt_radix8_kernel
// (include/ntt.cuh:30-62, one 32-lane warp per row of V polynomials) and the
// radix-2 128-thread form (ntt.cuh:308-328).  Here one wavefront transforms
// one polynomial and walks the batch with a grid-stride loop; the launch is
// sized to fill every CU (256 CUs x 8 XCDs on MI355X).
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "ntt.hpp"

namespace dil {

// wavefronts per workgroup: 4 by default, or 2 / 8 as a tiling chooses
constexpr int NTT_WAVES_DEFAULT = 4;

template <int W>
struct NttShared {
    ZetaLds zt;
    int32_t x[W][NTT_XBUF];
};

__device__ __forceinline__ void load_A(int32_t (&v)[4], const int32_t *p, int l) {
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = p[64 * r + l];
}
__device__ __forceinline__ void store_A(const int32_t (&v)[4], int32_t *p, int l) {
#pragma unroll
    for (int r = 0; r < 4; ++r) p[64 * r + l] = v[r];
}
__device__ __forceinline__ void load_D(int32_t (&v)[4], const int32_t *p, int l) {
    const int4 t = reinterpret_cast<const int4 *>(p)[l];
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
}
__device__ __forceinline__ void store_D(const int32_t (&v)[4], int32_t *p, int l) {
    reinterpret_cast<int4 *>(p)[l] = make_int4(v[0], v[1], v[2], v[3]);
}

// mode 0: forward in place; 1: inverse in place;
// 2: c = INTT(NTT(a) o_mont b) with b already in the NTT domain.
template <int KIND, int W>
__global__ __launch_bounds__(64 * W) void ntt_batch_kernel(int32_t *c, const int32_t *a, const int32_t *b,
                                                           size_t npoly, size_t pitch) {
    __shared__ NttShared<W> sh;
    zeta_lds_load(sh.zt);
    __syncthreads();
    const int l = lane_id();
    const int w = threadIdx.x >> 6;
    int32_t *x = sh.x[w];
    const size_t nwaves = (size_t)gridDim.x * W;
    for (size_t p = (size_t)blockIdx.x * W + w; p < npoly; p += nwaves) {
        int32_t v[4];
        if constexpr (KIND == 0) {
            load_A(v, a + p * pitch, l);
            ntt_wave(v, x, sh.zt);
            store_D(v, c + p * pitch, l);
        } else if constexpr (KIND == 1) {
            load_D(v, a + p * pitch, l);
            intt_wave(v, x, sh.zt);
            store_A(v, c + p * pitch, l);
        } else {
            load_A(v, a + p * pitch, l);
            ntt_wave(v, x, sh.zt);
            int32_t bb[4];
            load_D(bb, b + p * pitch, l);
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = mont_mul(v[r], bb[r]);
            intt_wave(v, x, sh.zt);
            store_A(v, c + p * pitch, l);
        }
    }
}

template <int W>
static int launch_ntt_w(int kind, int32_t *c, const int32_t *a, const int32_t *b, size_t npoly, size_t pitch_words,
                        int blocks, hipStream_t stream) {
    if (blocks <= 0) blocks = default_grid_blocks(npoly, W, 32 / W);
    dim3 grid(blocks), block(64 * W);
    switch (kind) {
        case 0:
            hipLaunchKernelGGL((ntt_batch_kernel<0, W>), grid, block, 0, stream, c, a, b, npoly, pitch_words);
            break;
        case 1:
            hipLaunchKernelGGL((ntt_batch_kernel<1, W>), grid, block, 0, stream, c, a, b, npoly, pitch_words);
            break;
        default:
            hipLaunchKernelGGL((ntt_batch_kernel<2, W>), grid, block, 0, stream, c, a, b, npoly, pitch_words);
            break;
    }
    return (int)hipGetLastError();
}

bool ntt_waves_supported(int w) { return w == 0 || w == 2 || w == 4 || w == 8; }

int launch_ntt(int kind, int32_t *c, const int32_t *a, const int32_t *b, size_t npoly, size_t pitch_words,
               int blocks, hipStream_t stream) {
    if (npoly == 0) return 0;
    switch (tiling_get(TILE_NTT_WAVES, 0, npoly)) {
        case 2: return launch_ntt_w<2>(kind, c, a, b, npoly, pitch_words, blocks, stream);
        case 8: return launch_ntt_w<8>(kind, c, a, b, npoly, pitch_words, blocks, stream);
        default: return launch_ntt_w<NTT_WAVES_DEFAULT>(kind, c, a, b, npoly, pitch_words, blocks, stream);
    }
}

// Keccak-f[1600] over n independent states (one per lane), for testing the
// permutation against the oracle.
__global__ void keccak_batch_kernel(uint64_t *st, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Lane64 s[25];
#pragma unroll
    for (int j = 0; j < 25; ++j) s[j] = to_lane(st[j * n + i]);
    keccak_f1600(s);
#pragma unroll
    for (int j = 0; j < 25; ++j) st[j * n + i] = from_lane(s[j]);
}

int launch_keccak(uint64_t *st, size_t n, hipStream_t stream) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(keccak_batch_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, st, n);
    return (int)hipGetLastError();
}

}  // namespace dil
