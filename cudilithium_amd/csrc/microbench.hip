// VALU throughput microbenchmark: lane-operations per second of the
// instruction classes the engine is built from, measured on the device so
// that roofline peaks are the card's own.
//
// Each lane runs 8 independent dependency chains of one operation for
// `iters` iterations; the launch fills every SIMD several waves deep.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "keccak.hpp"
#include "keccak_row.hpp"

namespace dil {

template <int OP>
__global__ __launch_bounds__(256) void valu_bench_kernel(uint32_t *out, uint32_t iters, uint32_t seed) {
    uint32_t x[8];
    double d[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        x[c] = seed * (threadIdx.x + 17 * c + 1) + blockIdx.x;
        d[c] = (double)x[c];
    }
    const uint32_t k = seed | 1u, k2 = seed ^ 0x9E3779B9u;
    const double dk = 1.0000001, db = 0.5;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if constexpr (OP == 0) x[c] = x[c] + k;                                    // v_add_u32
            if constexpr (OP == 1) x[c] = bitop3<0x96>(x[c], k, k2);                    // v_bitop3_b32
            if constexpr (OP == 2) x[c] = x[c] * k;                                     // v_mul_lo_u32
            if constexpr (OP == 3) x[c] = (uint32_t)__mulhi((int)x[c], (int)k);         // v_mul_hi_i32
            if constexpr (OP == 4) d[c] = fma(d[c], dk, db);                            // v_fma_f64
            if constexpr (OP == 5) x[c] = funnel(x[c], k, 13);                          // v_alignbit_b32
            if constexpr (OP == 6) x[c] = (uint32_t)(int)((double)(int)x[c] * dk) + k;  // cvt, mul_f64, cvt
            if constexpr (OP == 7) x[c] = __umul24(x[c], k);                             // v_mul_u32_u24
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) acc ^= x[c] ^ (uint32_t)(int64_t)d[c];
    if (acc == 0x12345678u) out[blockIdx.x] = acc;  // keeps the chains live
}

// Returns lane-operations per second of OP (0 add, 1 bitop3, 2 mul_lo,
// 3 mul_hi, 4 fma_f64, 5 alignbit, 6 cvt i32->f64->i32 pair, 7 mul_u24).
int valu_microbench(int op, uint32_t iters, double *ops_per_s) {
    const int cus = device_cu_count();
    const int blocks = cus * 16;  // 16 x 256 threads per CU: 16 waves per CU
    uint32_t *out = nullptr;
    hipError_t e = hipMalloc(&out, blocks * sizeof(uint32_t));
    if (e != hipSuccess) return (int)e;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    auto launch = [&](uint32_t it) {
        switch (op) {
            case 0: hipLaunchKernelGGL(valu_bench_kernel<0>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
            case 1: hipLaunchKernelGGL(valu_bench_kernel<1>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
            case 2: hipLaunchKernelGGL(valu_bench_kernel<2>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
            case 3: hipLaunchKernelGGL(valu_bench_kernel<3>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
            case 4: hipLaunchKernelGGL(valu_bench_kernel<4>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
            case 5: hipLaunchKernelGGL(valu_bench_kernel<5>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
            case 6: hipLaunchKernelGGL(valu_bench_kernel<6>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
            default: hipLaunchKernelGGL(valu_bench_kernel<7>, dim3(blocks), dim3(256), 0, 0, out, it, 3u); break;
        }
    };
    launch(16);  // warm-up
    (void)hipEventRecord(t0, 0);
    launch(iters);
    (void)hipEventRecord(t1, 0);
    e = hipEventSynchronize(t1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(out);
    if (e != hipSuccess) return (int)e;
    *ops_per_s = (double)blocks * 256.0 * 8.0 * iters / (ms * 1e-3);
    return (int)hipGetLastError();
}

// Keccak-f[1600] chains in the three state layouts: every thread runs
// `nperm` dependent permutations of its state (lane form: one state per lane;
// pair form: one per two lanes; row form: one per 16-lane row).
template <int FORM>
__global__ __launch_bounds__(256) void keccak_bench_kernel(uint32_t *out, uint32_t nperm) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    if constexpr (FORM == 0) {
        Lane64 s[25];
#pragma unroll
        for (int i = 0; i < 25; ++i) s[i] = {t * 31u + i, t ^ (i * 7u)};
        for (uint32_t p = 0; p < nperm; ++p) keccak_f1600(s);
#pragma unroll
        for (int i = 0; i < 25; ++i) acc ^= s[i].lo ^ s[i].hi;
    } else if constexpr (FORM == 1) {
        uint32_t s[25];
#pragma unroll
        for (int i = 0; i < 25; ++i) s[i] = t * 31u + i;
        for (uint32_t p = 0; p < nperm; ++p) keccak_f1600_pair(s, t & 1);
#pragma unroll
        for (int i = 0; i < 25; ++i) acc ^= s[i];
    } else {
        __shared__ uint32_t lds[4 * 4 * ROW_LDS_WORDS];
        const RowLane r = row_lane_setup(lds, (threadIdx.x >> 6) * 4 * ROW_LDS_WORDS);
        uint32_t s[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) s[i] = r.active ? t * 31u + i : 0u;
        for (uint32_t p = 0; p < nperm; ++p) keccak_f1600_row(s, r, lds);
        acc = s[0] ^ s[1] ^ s[2] ^ s[3] ^ s[4];
    }
    out[t] = acc;
}

// form 0 lane, 1 pair, 2 row; blocks of 256 threads (threads = 64: one wave).
// Returns microseconds per chained permutation and permutations per second.
int keccak_microbench(int form, uint32_t blocks, uint32_t threads, uint32_t nperm, double *us_per_perm,
                      double *perms_per_s) {
    if (form < 0 || form > 2 || blocks == 0 || (threads != 64 && threads != 256) || nperm == 0) return -1;
    uint32_t *out = nullptr;
    hipError_t e = hipMalloc(&out, (size_t)blocks * threads * sizeof(uint32_t));
    if (e != hipSuccess) return (int)e;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    auto launch = [&](uint32_t np) {
        if (form == 0) hipLaunchKernelGGL(keccak_bench_kernel<0>, dim3(blocks), dim3(threads), 0, 0, out, np);
        if (form == 1) hipLaunchKernelGGL(keccak_bench_kernel<1>, dim3(blocks), dim3(threads), 0, 0, out, np);
        if (form == 2) hipLaunchKernelGGL(keccak_bench_kernel<2>, dim3(blocks), dim3(threads), 0, 0, out, np);
    };
    launch(4);  // warm-up
    (void)hipEventRecord(t0, 0);
    launch(nperm);
    (void)hipEventRecord(t1, 0);
    e = hipEventSynchronize(t1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(out);
    if (e != hipSuccess) return (int)e;
    const double states_per_thread = form == 0 ? 1.0 : form == 1 ? 0.5 : 1.0 / 16.0;
    *us_per_perm = ms * 1e3 / nperm;
    *perms_per_s = (double)blocks * threads * states_per_thread * nperm / (ms * 1e-3);
    return (int)hipGetLastError();
}

}  // namespace dil
