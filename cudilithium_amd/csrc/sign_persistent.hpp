// Persistent work-queue signing engine.
//
// Replaces the reference's host rejection scheduler (api.cu:208-441: one
// stream synchronisation and ~8 host<->device transfers per round) with one
// launch in which every workgroup loops until the batch is done:
//
//   refill   claim items from a global queue (one atomicAdd per refill) until
//            the workgroup owns `claim` (at most S) undone items or the queue
//            is empty; with guided claims fewer near the end of the queue
//   assign   S attempt slots round-robin over the owned items: item o gets
//            slots o, o + U, o + 2U, ... with consecutive nonces, so freed
//            slots turn into speculative nonces of unfinished items instead of
//            idling at the end of the batch
//   y w c    the first three attempt stages on the S slots (sign_stages.hpp),
//            with the lane mapping each stage wants, separated by barriers
//   z        one wave per item: its slots in slot order = nonce order; the
//            first that passes is the smallest successful nonce -- the
//            sequential specification's result -- and writes the signature
//            (z, h and c~ straight into the item's signature row)
//
// Workgroups never wait on each other (the queue is only read-modify-written
// by atomicAdd), so there is no co-residency requirement and no deadlock: a
// workgroup that starts after the queue has drained exits at once.  Every
// round gives each owned item at least one attempt, so each workgroup
// finishes with probability 1, as the reference's loop does.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"
#include "sign_stages.hpp"

namespace dil {

// threads per workgroup (4 waves; a build may choose 2 with -DDIL_PERSIST_THREADS=128,
// then with 16 slots per workgroup so that the y stage's S * L sponges fit one pass)
#ifndef DIL_PERSIST_THREADS
#define DIL_PERSIST_THREADS 256
#endif
constexpr int PERSIST_THREADS = DIL_PERSIST_THREADS;
constexpr uint32_t PERSIST_MAX_ROUNDS = 20000;
// control words (uint32): 0 queue head, 1 status, 2 attempts executed.
// Optional phase profile: PERSIST_PHASES uint64 cycle sums (assign, y, w, c,
// z, commit) at a separate pointer, null when not profiling.
constexpr int PERSIST_PHASES = 6;
// wave priority (s_setprio) of the Keccak stages; 0 elsewhere.  Measured on
// 16384-item batches: +2-3 % signatures/s for Dilithium3 over no priority.
#ifndef DIL_KECCAK_PRIO
#define DIL_KECCAK_PRIO 2
#endif
constexpr int KECCAK_PRIO = DIL_KECCAK_PRIO;
// claim argument of the guided kernel: items per workgroup (low 16 bits) and
// the number of workgroups resident at once (high 16 bits); guided claims
// are built for the 2-waves-per-SIMD register budget only
constexpr int CLAIM_GUIDED_FLAG = 0x10000;  // DIL_CLAIM_GUIDED of the header

// WPE: waves per SIMD the register allocation must allow (512 / WPE VGPRs)
// claim: at most this many items owned per workgroup (1..S); below S every
// owned item gets at least S / claim slots (nonces) per round, which trades
// speculative attempts for fewer rounds on small batches
// GUIDED: guided claims (a separate instantiation, so that the scalar code
// for it does not change the register allocation of the plain kernel)
template <class P, int S, int WPE, bool GUIDED>
__global__ __launch_bounds__(PERSIST_THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void sg_persistent_kernel(
    SignArgs a, SlotRef sr, uint32_t *head, int claim, unsigned long long *phase_sums) {
    __shared__ ZetaLds zt;
    __shared__ int32_t x[PERSIST_THREADS / 64][NTT_XBUF];
    // L = 7: the paired w stage's transformed y vectors (56 KiB per workgroup,
    // which leaves room for two workgroups per CU: with more waves per SIMD
    // the w stage does not pair for L = 7)
    constexpr bool W_Y_LDS = P::L >= DIL_W_Y_LDS_MIN_L && WPE <= 2;
    constexpr bool W_PAIRS = P::L < DIL_W_Y_LDS_MIN_L || W_Y_LDS;
    __shared__ int32_t ylds[W_Y_LDS ? PERSIST_THREADS / 64 : 1][W_Y_LDS ? 2 * P::L * N : 1];
    __shared__ uint32_t own_item[S], own_next[S];
    __shared__ int own_done[S];
    __shared__ uint32_t slot_own[S], slot_nonce[S];
    __shared__ uint8_t w_unit[S];  // w work units: slot s, paired with s + u when s / u is even
    __shared__ int n_own, queue_empty, n_units;
    const int tid = threadIdx.x, wave = tid >> 6, lane = lane_id();
    constexpr int NW = PERSIST_THREADS / 64;
    const size_t slot0 = (size_t)blockIdx.x * S;
    auto slot_base = [&](int s) { return sr.base + (slot0 + s) * sr.pitch; };
    auto slot_sig = [&](int s) { return sr.sig + (slot0 + s) * sr.sig_pitch; };
    zeta_lds_load(zt);
    if (tid == 0) {
        n_own = 0;
        queue_empty = 0;
    }
    __syncthreads();
    // optional phase profile: workgroup cycles per phase, summed over
    // workgroups; kept in LDS (only thread 0 uses it) so that it holds no
    // registers through the loop
    __shared__ unsigned long long tph[PERSIST_PHASES + 1];  // [PERSIST_PHASES]: last mark
    const bool prof = phase_sums != nullptr;
    if (prof && tid == 0) {
        for (int p = 0; p < PERSIST_PHASES; ++p) tph[p] = 0;
        tph[PERSIST_PHASES] = clock64();
    }
    auto mark = [&](int ph) {
        if (prof && tid == 0) {
            const unsigned long long t = clock64();
            tph[ph] += t - tph[PERSIST_PHASES];
            tph[PERSIST_PHASES] = t;
        }
    };

    for (uint32_t round = 0;; ++round) {
        if (round == PERSIST_MAX_ROUNDS) {
            // no correct run gets here (each round succeeds per item with
            // probability >= 1/5); stop instead of spinning and report it
            if (tid == 0) atomicOr(head + 1, SIGN_STATUS_ROUND_CAP);
            break;
        }
        if (tid == 0) {
            const int claim_items = GUIDED ? claim & 0xFFFF : claim;
            int cap = claim_items > 0 && claim_items < S ? claim_items : S;
            if constexpr (GUIDED) {
                // guided claims: once fewer than cap items per resident
                // workgroup are left in the queue, own at most remaining /
                // resident workgroups of them (rounded up), so that the last
                // items spread over all workgroups (each then gets more slots
                // per round) instead of holding a few workgroups while the
                // others exit
                const uint32_t resident = (uint32_t)claim >> 16;  // >= 1: the launch passes it
                if (!queue_empty) {
                    const uint32_t h0 = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const size_t rem = (size_t)h0 < a.n ? a.n - h0 : 0;
                    if (rem < (size_t)cap * resident) {
                        int per = 1;
                        while ((size_t)per * resident < rem) ++per;
                        cap = per;
                    }
                }
            }
            int u = n_own;
            if (u < cap && !queue_empty) {
                const uint32_t start = atomicAdd(head, (uint32_t)(cap - u));
                for (int j = 0; j < cap - u && (size_t)start + j < a.n; ++j) {
                    own_item[n_own] = start + j;
                    own_next[n_own] = 0;
                    ++n_own;
                }
                if ((size_t)start + (cap - u) >= a.n) queue_empty = 1;
            }
            u = n_own;
            for (int s = 0; s < S && u > 0; ++s) {
                const int o = s % u;
                slot_own[s] = o;
                slot_nonce[s] = own_next[o] + s / u;
            }
            for (int o = 0; o < u; ++o) {
                own_next[o] += (S - o + u - 1) / u;
                own_done[o] = 0;
            }
            // slots s and s + u carry consecutive nonces of the same item: the
            // w stage takes them as one unit (one pass over that item's A)
            int nu = 0;
            for (int s = 0; s < S && u > 0; ++s)
                if ((s / u) % 2 == 0) w_unit[nu++] = (uint8_t)s;
            n_units = nu;
        }
        __syncthreads();
        mark(0);
        const int u = n_own;
        if (u == 0) break;
        if (tid == 0) atomicAdd(head + 2, (uint32_t)S);  // attempts executed (incl. speculative)

        // the Keccak stages (y, c) run few serial permutation chains that the
        // whole workgroup then waits for; their waves get issue priority over
        // the transform stages of other workgroups on the same SIMDs
        __builtin_amdgcn_s_setprio(KECCAK_PRIO);
        // When the S * L sponges fit two lanes each, the pair form's shorter
        // chain (Dilithium2 16,384 items 2.41 -> 2.28 ms per batch; 16 slots per
        // workgroup: Dilithium3 / 5 4,096 items 1.79 / 1.73 -> 1.67 / 1.67 ms)
#ifndef DIL_Y_PAIR
#define DIL_Y_PAIR 1
#endif
        if constexpr (DIL_Y_PAIR && 2 * S * P::L <= PERSIST_THREADS) {
            if (tid < 2 * S * P::L) {
                const int q = tid >> 1, s = q / P::L;
                stage_y_pair<P>(a, own_item[slot_own[s]], slot_nonce[s], q % P::L, slot_base(s), tid & 1);
            }
        } else {
            for (int t = tid; t < S * P::L; t += PERSIST_THREADS) {
                const int s = t / P::L;
                stage_y<P>(a, own_item[slot_own[s]], slot_nonce[s], t % P::L, slot_base(s));
            }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        mark(1);
        for (int t = wave; t < n_units; t += NW) {
            const int s = w_unit[t];
            const size_t item = own_item[slot_own[s]];
            if (s + u < S) {
                if constexpr (W_PAIRS) {
                    uint8_t *const pair[2] = {slot_base(s), slot_base(s + u)};
                    stage_w_multi<P, 2>(a, item, pair, x[wave], zt, ylds[W_Y_LDS ? wave : 0]);
                    continue;
                } else {
                    stage_w<P>(a, item, slot_base(s + u), x[wave], zt);
                }
            }
            stage_w<P>(a, item, slot_base(s), x[wave], zt);
        }
        __syncthreads();
        mark(2);
        __builtin_amdgcn_s_setprio(KECCAK_PRIO);
        // two lanes per slot (keccak_f1600_pair): S slots on 2 S lanes
        for (int t = tid; t < 2 * S; t += PERSIST_THREADS) {
            const int s = t >> 1;
            stage_c_pair<P>(a, own_item[slot_own[s]], slot_base(s), slot_sig(s), (uint32_t)(t & 1));
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        mark(3);
        // z and commit, one wave per owned item: its slots o, o + u, ... carry
        // increasing nonces, so they are checked in that order and the first
        // that passes -- the sequential specification's result -- ends the
        // item; later (speculative) slots of the item are skipped.  An attempt
        // writes z and h straight into the item's signature; a rejected one
        // leaves them partly written, and the accepted attempt rewrites every
        // byte (c~ is copied from the slot where the c stage put it).
        for (int o = wave; o < u; o += NW) {
            const size_t item = own_item[o];
            uint8_t *dst = a.sig + item * a.sig_pitch;
            for (int s = o; s < S; s += u) {
                if (stage_z<P>(a, item, slot_base(s), dst, x[wave], zt)) {
                    const uint8_t *src = slot_sig(s);
                    if (lane < SEEDBYTES) dst[lane] = src[lane];
                    if (lane == 0) {
                        own_done[o] = 1;
                        // attempts the sequential specification makes for this item
                        *reinterpret_cast<uint32_t *>(a.rows + item * a.row_pitch + SignLayout<P>::ATTEMPTS) =
                            slot_nonce[s] + 1;
                    }
                    break;
                }
            }
        }
        __syncthreads();
        mark(4);
        if (tid == 0) {
            int k = 0;
            for (int o = 0; o < u; ++o)
                if (!own_done[o]) {
                    own_item[k] = own_item[o];
                    own_next[k] = own_next[o];
                    ++k;
                }
            n_own = k;
        }
        __syncthreads();
        mark(5);
    }
    if (prof && tid == 0) {
        for (int p = 0; p < PERSIST_PHASES; ++p) atomicAdd(phase_sums + p, tph[p]);
    }
}

// clears the control words (and the phase sums when given); a kernel rather
// than a memset so that the sequence can be captured in a HIP graph
__global__ void zero_ctl_kernel(uint32_t *p, unsigned long long *phase_sums);

template <class P, int S, int WPE>
static int launch_persistent_s(const SignArgs &a, const SlotRef &sr, size_t nslots, int blocks, int claim,
                               uint32_t *d_head, unsigned long long *phase_sums, hipStream_t st) {
    int nb = (int)(nslots / S);
    if (blocks > 0 && blocks < nb) nb = blocks;
    if (nb < 1) return -3;
    hipLaunchKernelGGL(zero_ctl_kernel, dim3(1), dim3(1), 0, st, d_head, phase_sums);
    // claim: items per workgroup (low 16 bits) | workgroups resident at once
    // (high 16 bits: the register budget allows WPE waves per SIMD)
    if ((claim & CLAIM_GUIDED_FLAG) && WPE == 2) {
        const int resident = std::max(1, std::min(nb, device_cu_count() * 4 * WPE / (PERSIST_THREADS / 64)));
        const int claim_arg = (claim & 0xFFFF) | (std::min(resident, 0xFFFF) << 16);
        hipLaunchKernelGGL((sg_persistent_kernel<P, S, WPE, WPE == 2>), dim3(nb), dim3(PERSIST_THREADS), 0, st, a,
                           sr, d_head, claim_arg, phase_sums);
    } else {
        hipLaunchKernelGGL((sg_persistent_kernel<P, S, WPE, false>), dim3(nb), dim3(PERSIST_THREADS), 0, st, a, sr,
                           d_head, claim & 0xFFFF, phase_sums);
    }
    return (int)hipGetLastError();
}

// Defined once per mode (sign_persistent_<mode>.hip).  occupancy: waves per
// SIMD the kernel's register allocation allows (2, 3 or 4).
template <class P>
int launch_persistent_p(const SignArgs &a, const SlotRef &sr, size_t nslots, int slots_per_block, int occupancy,
                        int blocks, int claim, uint32_t *d_head, unsigned long long *phase_sums, hipStream_t st);

template <class P, int WPE>
static int launch_persistent_w(const SignArgs &a, const SlotRef &sr, size_t nslots, int slots_per_block, int blocks,
                               int claim, uint32_t *d_head, unsigned long long *ph, hipStream_t st) {
    switch (slots_per_block) {
        case 1: return launch_persistent_s<P, 1, WPE>(a, sr, nslots, blocks, claim, d_head, ph, st);
        case 8: return launch_persistent_s<P, 8, WPE>(a, sr, nslots, blocks, claim, d_head, ph, st);
        case 16: return launch_persistent_s<P, 16, WPE>(a, sr, nslots, blocks, claim, d_head, ph, st);
        case 32: return launch_persistent_s<P, 32, WPE>(a, sr, nslots, blocks, claim, d_head, ph, st);
        default: return -1;
    }
}

#define DIL_DEFINE_PERSISTENT_LAUNCHER(MODE)                                                                       \
    template <>                                                                                                  \
    int launch_persistent_p<Params<MODE>>(const SignArgs &a, const SlotRef &sr, size_t nslots,                    \
                                          int slots_per_block, int occupancy, int blocks, int claim,              \
                                          uint32_t *d_head, unsigned long long *ph, hipStream_t st) {            \
        using P = Params<MODE>;                                                                                  \
        switch (occupancy) {                                                                                     \
            case 2: return launch_persistent_w<P, 2>(a, sr, nslots, slots_per_block, blocks, claim, d_head, ph, st); \
            case 3: return launch_persistent_w<P, 3>(a, sr, nslots, slots_per_block, blocks, claim, d_head, ph, st); \
            case 4: return launch_persistent_w<P, 4>(a, sr, nslots, slots_per_block, blocks, claim, d_head, ph, st); \
            default: return -1;                                                                                  \
        }                                                                                                        \
    }

}  // namespace dil
