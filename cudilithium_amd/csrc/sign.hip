// Batched deterministic signing: kernels of the per-stage form and its
// host-driven rejection scheduler.
//
// The per-stage form restates the reference architecture (api.cu:63-452):
// one launch per stage, attempts issued by the host in rounds of
// exec_threshold (item << 16 | nonce) entries, a stream synchronisation per
// round, full rounds written in place and tail rounds issued with several
// consecutive nonces per item into temporary rows, the smallest successful
// nonce per item winning (api.cu:218-441).  It is the baseline the
// persistent engine (sign_persistent.hip) is measured against.
#include <hip/hip_runtime.h>

#include <vector>

#include "internal.hpp"
#include "sign_stages.hpp"

namespace dil {

constexpr int SIGN_WAVES = 4;

// One lane absorbs a whole message serially, so the batch waits for its
// longest message; the wave asks for the highest issue priority on its SIMD
// (s_setprio, a per-wave hint) so that other kernels' waves sharing the SIMD
// do not stretch that serial chain.
template <class P>
__global__ __launch_bounds__(256) void sg_mu_kernel(SignArgs a) {
    __builtin_amdgcn_s_setprio(3);
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) stage_mu<P>(a, i);
}

// two lanes per item: 128 items per workgroup
template <class P>
__global__ __launch_bounds__(256) void sg_mu_pair_kernel(SignArgs a) {
    __builtin_amdgcn_s_setprio(3);
    const size_t i = (size_t)blockIdx.x * 128 + (threadIdx.x >> 1);
    if (i < a.n) stage_mu_pair<P>(a, i, threadIdx.x & 1);  // both lanes of a pair take the same branch
}

template <class P>
__global__ __launch_bounds__(256) void sg_expand_kernel(SignArgs a) {
    constexpr int KL = P::K * P::L;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * KL) return;
    const int e = (int)(t % KL);
    stage_expand_a<P>(a, t / KL, e / P::L, e % P::L);
}

template <class P>
__global__ __launch_bounds__(64 * SIGN_WAVES) void sg_prep_kernel(SignArgs a) {
    constexpr int NP = P::L + 2 * P::K;
    __shared__ ZetaLds zt;
    __shared__ int32_t x[SIGN_WAVES][NTT_XBUF];
    zeta_lds_load(zt);
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const size_t total = a.n * NP, stride = (size_t)gridDim.x * SIGN_WAVES;
    for (size_t t = (size_t)blockIdx.x * SIGN_WAVES + w; t < total; t += stride)
        stage_prep<P>(a, t / NP, (int)(t % NP), x[w], zt);
}

// exec entry e = (item << 16) | nonce (api.cu:221-230)
template <class P>
__global__ __launch_bounds__(256) void sg_y_kernel(SignArgs a, SlotRef sr, const uint32_t *exec, size_t nslot,
                                                   int by_item) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nslot * P::L) return;
    const size_t s = t / P::L;
    const int l = (int)(t % P::L);
    const uint32_t e = exec[s];
    const size_t item = e >> 16;
    stage_y<P>(a, item, e & 0xFFFF, l, sr.base + (by_item ? item : s) * sr.pitch);
}

template <class P>
__global__ __launch_bounds__(64 * SIGN_WAVES) void sg_w_kernel(SignArgs a, SlotRef sr, const uint32_t *exec,
                                                               size_t nslot, int by_item) {
    __shared__ ZetaLds zt;
    __shared__ int32_t x[SIGN_WAVES][NTT_XBUF];
    zeta_lds_load(zt);
    __syncthreads();
    const int w = threadIdx.x >> 6;
    for (size_t s = (size_t)blockIdx.x * SIGN_WAVES + w; s < nslot; s += (size_t)gridDim.x * SIGN_WAVES) {
        const size_t item = exec[s] >> 16;
        stage_w<P>(a, item, sr.base + (by_item ? item : s) * sr.pitch, x[w], zt);
    }
}

template <class P>
__global__ __launch_bounds__(256) void sg_c_kernel(SignArgs a, SlotRef sr, const uint32_t *exec, size_t nslot,
                                                   int by_item) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const size_t item = exec[s] >> 16;
    const size_t r = by_item ? item : s;
    stage_c<P>(a, item, sr.base + r * sr.pitch, sr.sig + r * sr.sig_pitch);
}

template <class P>
__global__ __launch_bounds__(64 * SIGN_WAVES) void sg_z_kernel(SignArgs a, SlotRef sr, const uint32_t *exec,
                                                               size_t nslot, int by_item, uint8_t *done) {
    __shared__ ZetaLds zt;
    __shared__ int32_t x[SIGN_WAVES][NTT_XBUF];
    zeta_lds_load(zt);
    __syncthreads();
    const int w = threadIdx.x >> 6;
    for (size_t s = (size_t)blockIdx.x * SIGN_WAVES + w; s < nslot; s += (size_t)gridDim.x * SIGN_WAVES) {
        const size_t item = exec[s] >> 16;
        const size_t r = by_item ? item : s;
        const bool ok = stage_z<P>(a, item, sr.base + r * sr.pitch, sr.sig + r * sr.sig_pitch, x[w], zt);
        if (lane_id() == 0) done[s] = ok ? 1 : 0;
    }
}

// copies the winning temporary signature of each listed item; entries are
// {exec_idx, sign_idx} pairs as in copy_lut_element (api.cuh:36-39,
// sign.cu:1198-1207)
template <class P>
__global__ __launch_bounds__(256) void sg_copy_kernel(SignArgs a, SlotRef sr, const uint64_t *copy, size_t ncopy) {
    using Z = Sizes<P>;
    for (size_t c = blockIdx.x; c < ncopy; c += gridDim.x) {
        const uint8_t *src = sr.sig + copy[2 * c] * sr.sig_pitch;
        uint8_t *dst = a.sig + copy[2 * c + 1] * a.sig_pitch;
        for (int b = threadIdx.x; b < Z::SIG_BYTES; b += blockDim.x) dst[b] = src[b];
    }
}

// Resident-key signing: an item whose key was loaded without its secret key
// (FLAGS lacks KEY_HAS_SECRET; its secret arrays are zero) gets an all-zero
// signature and sets SIGN_STATUS_NO_SECRET in *status.  Runs after the
// rejection loop, lane per item.
template <class P>
__global__ __launch_bounds__(256) void sg_key_check_kernel(SignArgs a, uint32_t *status) {
    using KY = KeyLayout<P>;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint8_t *key_row = a.key_ints - KY::MAT + sk_index(a, i) * a.key_pitch;
    if (*reinterpret_cast<const uint32_t *>(key_row + KY::FLAGS) & KEY_HAS_SECRET) return;
    uint8_t *sig = a.sig + i * a.sig_pitch;
    for (int b = 0; b < Sizes<P>::SIG_BYTES; ++b) sig[b] = 0;
    atomicOr(status, SIGN_STATUS_NO_SECRET);
}

template <class P>
int sign_key_check_impl(const SignArgs &a, uint32_t *status, hipStream_t st) {
    if (a.n == 0 || !a.key_idx) return 0;
    hipLaunchKernelGGL(sg_key_check_kernel<P>, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a, status);
    return (int)hipGetLastError();
}

// The three fixed-stage launches are independent: mu/rhoprime (two lanes per
// item, latency-bound on long messages), ExpandA (one lane per matrix entry,
// throughput-bound) and the s1/s2/t0 NTTs run as a fork/join over the
// caller's stream and two auxiliary streams of this device, so the idle
// lanes of one fill with the others' waves.  Event fork/join is also what a
// HIP graph capture records as parallel branches.
template <class P>
int sign_fixed_impl(const SignArgs &a, hipStream_t st) {
    if (a.n == 0) return 0;
    constexpr int KL = P::K * P::L, NP = P::L + 2 * P::K;
    auto mu = [&](hipStream_t s) -> int {
        if (a.row_mu) return launch_sign_mu_row(P::MODE, a, s);
        // two lanes per item by default: fixed stage of 16,384 Dilithium3
        // items 0.77 -> 0.53 ms when launched call by call; inside a graph,
        // where ExpandA and the transforms bound it, the step is unchanged
#ifndef DIL_MU_PAIR
#define DIL_MU_PAIR 1
#endif
        if (DIL_MU_PAIR)
            hipLaunchKernelGGL(sg_mu_pair_kernel<P>, dim3((unsigned)((a.n + 127) / 128)), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(sg_mu_kernel<P>, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
        return (int)hipGetLastError();
    };
    if (a.key_idx) return mu(st);  // resident keys carry A, s1^, s2^, t0^ already: only mu / rhoprime
    // Batches below DIL_FIXED_FORK_MIN items run the three launches in order
    // on the caller's stream.  Off by default: for one caller the fork is
    // faster (1,024 / 4,096 items: 1.76 / 1.83 ms against 1.86 / 2.07 ms);
    // with ten concurrent callers the serial form was faster (4.6 against
    // 5.0 ms per round of 1,000-item calls).
#ifndef DIL_FIXED_FORK_MIN
#define DIL_FIXED_FORK_MIN 0
#endif
    if (a.n < (size_t)DIL_FIXED_FORK_MIN) {
        int e;
        if ((e = mu(st))) return e;
        hipLaunchKernelGGL(sg_prep_kernel<P>, dim3(default_grid_blocks(a.n * NP, SIGN_WAVES)), dim3(64 * SIGN_WAVES),
                           0, st, a);
        hipLaunchKernelGGL(sg_expand_kernel<P>, dim3((unsigned)((a.n * KL + 255) / 256)), dim3(256), 0, st, a);
        return (int)hipGetLastError();
    }
    AuxStreams *aux = aux_streams(st);
    if (!aux) return (int)hipErrorOutOfMemory;
    std::lock_guard<std::mutex> guard(aux->lock);
    int e;
    if ((e = (int)hipEventRecord(aux->fork, st))) return e;
    if ((e = (int)hipStreamWaitEvent(aux->s[0], aux->fork, 0))) return e;
    if ((e = (int)hipStreamWaitEvent(aux->s[1], aux->fork, 0))) return e;
    if ((e = mu(aux->s[0]))) return e;
    hipLaunchKernelGGL(sg_prep_kernel<P>, dim3(default_grid_blocks(a.n * NP, SIGN_WAVES)), dim3(64 * SIGN_WAVES), 0,
                       aux->s[1], a);
    hipLaunchKernelGGL(sg_expand_kernel<P>, dim3((unsigned)((a.n * KL + 255) / 256)), dim3(256), 0, st, a);
    if ((e = (int)hipGetLastError())) return e;
    if ((e = (int)hipEventRecord(aux->join[0], aux->s[0]))) return e;
    if ((e = (int)hipEventRecord(aux->join[1], aux->s[1]))) return e;
    if ((e = (int)hipStreamWaitEvent(st, aux->join[0], 0))) return e;
    return (int)hipStreamWaitEvent(st, aux->join[1], 0);
}

template <class P>
int sign_round_impl(const SignArgs &a, const SlotRef &sr, const uint32_t *d_exec, size_t nslot, int by_item,
                    uint8_t *d_done, hipStream_t st) {
    if (nslot == 0) return 0;
    const unsigned bl = (unsigned)((nslot * P::L + 255) / 256), bs = (unsigned)((nslot + 255) / 256);
    const int bw = default_grid_blocks(nslot, SIGN_WAVES);
    hipLaunchKernelGGL(sg_y_kernel<P>, dim3(bl), dim3(256), 0, st, a, sr, d_exec, nslot, by_item);
    hipLaunchKernelGGL(sg_w_kernel<P>, dim3(bw), dim3(64 * SIGN_WAVES), 0, st, a, sr, d_exec, nslot, by_item);
    hipLaunchKernelGGL(sg_c_kernel<P>, dim3(bs), dim3(256), 0, st, a, sr, d_exec, nslot, by_item);
    hipLaunchKernelGGL(sg_z_kernel<P>, dim3(bw), dim3(64 * SIGN_WAVES), 0, st, a, sr, d_exec, nslot, by_item, d_done);
    return (int)hipGetLastError();
}

template <class P>
int sign_copy_impl(const SignArgs &a, const SlotRef &sr, const uint64_t *d_copy, size_t ncopy, hipStream_t st) {
    if (ncopy == 0) return 0;
    const unsigned b = (unsigned)(ncopy < 4096 ? ncopy : 4096);
    hipLaunchKernelGGL(sg_copy_kernel<P>, dim3(b), dim3(256), 0, st, a, sr, d_copy, ncopy);
    return (int)hipGetLastError();
}

#define DIL_DISPATCH(mode, fn, ...)                      \
    switch (mode) {                                      \
        case 2: return fn<Params<2>>(__VA_ARGS__);       \
        case 3: return fn<Params<3>>(__VA_ARGS__);       \
        case 5: return fn<Params<5>>(__VA_ARGS__);       \
        default: return -1;                              \
    }

int launch_sign_fixed(int mode, const SignArgs &a, hipStream_t st) { DIL_DISPATCH(mode, sign_fixed_impl, a, st) }

int launch_sign_key_check(int mode, const SignArgs &a, uint32_t *status, hipStream_t st) {
    DIL_DISPATCH(mode, sign_key_check_impl, a, status, st)
}

int launch_sign_round(int mode, const SignArgs &a, const SlotRef &sr, const uint32_t *d_exec, size_t nslot,
                      int by_item, uint8_t *d_done, hipStream_t st) {
    DIL_DISPATCH(mode, sign_round_impl, a, sr, d_exec, nslot, by_item, d_done, st)
}

int launch_sign_copy(int mode, const SignArgs &a, const SlotRef &sr, const uint64_t *d_copy, size_t ncopy,
                     hipStream_t st) {
    DIL_DISPATCH(mode, sign_copy_impl, a, sr, d_copy, ncopy, st)
}

// Host rejection scheduler of the per-stage form (api.cu:208-441).  The
// caller's lookup tables are used exactly as the reference uses them:
//   sign_lut[n] {done_flag, nonce}        per-item host state
//   h/d_exec[thr]  (item << 16 | nonce)   attempts of one round
//   h/d_done[thr]                         success flag per attempt
//   h/d_copy[2*thr] {exec_idx, sign_idx}  winners of a tail round
int sign_perstage(int mode, const SignArgs &a, const SlotRef &item_slots, const SlotRef &temp_slots, size_t thr,
                  SignLutEntry *sign_lut, uint32_t *h_exec, uint32_t *d_exec, uint8_t *h_done, uint8_t *d_done,
                  uint64_t *h_copy, uint64_t *d_copy, hipStream_t st) {
    int e = launch_sign_fixed(mode, a, st);
    if (e) return e;
    for (size_t i = 0; i < a.n; ++i) sign_lut[i] = SignLutEntry{0, 0};
    size_t remaining = a.n;
    auto issue = [&](size_t &ne) {
        for (size_t i = 0; i < a.n && ne < thr; ++i)
            if (!sign_lut[i].done_flag) h_exec[ne++] = (uint32_t)(i << 16) | (sign_lut[i].nonce++ & 0xFFFFu);
    };
    while (remaining > 0) {
        size_t ne = 0;
        issue(ne);
        const bool full = ne == thr;
        while (ne < thr) issue(ne);  // tail: consecutive nonces round-robin
        if ((e = (int)hipMemcpyAsync(d_exec, h_exec, ne * sizeof(uint32_t), hipMemcpyHostToDevice, st))) return e;
        if ((e = (int)hipMemsetAsync(d_done, 0, ne, st))) return e;
        if ((e = launch_sign_round(mode, a, full ? item_slots : temp_slots, d_exec, ne, full ? 1 : 0, d_done, st)))
            return e;
        if ((e = (int)hipMemcpyAsync(h_done, d_done, ne, hipMemcpyDeviceToHost, st))) return e;
        if ((e = (int)hipStreamSynchronize(st))) return e;
        if (full) {
            for (size_t s = 0; s < ne; ++s)
                if (h_done[s]) {
                    sign_lut[h_exec[s] >> 16].done_flag = 1;
                    --remaining;
                }
        } else {
            // the first success of an item in issue order has its smallest nonce
            size_t nc = 0;
            for (size_t s = 0; s < ne; ++s) {
                const size_t item = h_exec[s] >> 16;
                if (h_done[s] && !sign_lut[item].done_flag) {
                    sign_lut[item].done_flag = 1;
                    --remaining;
                    h_copy[2 * nc] = s;
                    h_copy[2 * nc + 1] = item;
                    ++nc;
                }
            }
            if (nc) {
                if ((e = (int)hipMemcpyAsync(d_copy, h_copy, 2 * nc * sizeof(uint64_t), hipMemcpyHostToDevice, st)))
                    return e;
                if ((e = launch_sign_copy(mode, a, temp_slots, d_copy, nc, st))) return e;
            }
        }
    }
    return 0;
}

}  // namespace dil
