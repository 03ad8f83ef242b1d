// Argument blocks and launchers of the keygen / sign / verify pipelines.
// Every array is addressed as base + item * pitch so the same kernels serve
// the drop-in pitched pools and the device-resident batch API.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dil {

struct KeygenArgs {
    uint8_t *pk;
    size_t pk_pitch;
    uint8_t *sk;
    size_t sk_pitch;
    uint8_t *scratch;  // KeypairLayout::SCRATCH area of each row
    size_t scratch_pitch;
    const uint8_t *seeds;  // 32 bytes per key, or null for counter seeds
    size_t seed_pitch;
    uint64_t rand_index;
    size_t n;
};

// Messages: item i is msg + (moff ? moff[i] : i * mpitch), length
// (mlen ? mlen[i] : mlen_const).
struct MsgRef {
    const uint8_t *msg;
    const uint64_t *moff;
    size_t mpitch;
    const uint32_t *mlen;
    uint32_t mlen_const;
    __host__ __device__ const uint8_t *ptr(size_t i) const { return msg + (moff ? moff[i] : i * mpitch); }
    __host__ __device__ uint32_t len(size_t i) const { return mlen ? mlen[i] : mlen_const; }
};

struct VerifyArgs {
    const uint8_t *sig;
    size_t sig_pitch;
    const uint8_t *pk;
    size_t pk_pitch;
    MsgRef m;
    uint8_t *work;  // VerifyLayout rows (only HASH.. fields used)
    size_t work_pitch;
    int32_t *ret;
    size_t ret_pitch;  // in int32 units
    size_t n;
    // Resident keys (optional): item i verifies under key key_idx[i] of the
    // store at keys (nkeys KeyLayout rows of key_pitch bytes); pk / pk_pitch
    // are then ignored and tr, A and t1^ come from the store.  An index
    // >= nkeys rejects the item (ret -1) without touching the store.
    const uint32_t *key_idx;
    const uint8_t *keys;
    size_t key_pitch;
    uint32_t nkeys;
    // Long messages: tr and mu are computed before vf_prep with one sponge
    // per 16-lane row (vf_mu_row_kernel) and vf_prep skips them.
    int row_mu;
    // Otherwise, fork_mu: tr and mu are computed with two lanes per item on
    // an auxiliary stream beside vf_prep / vf_main (only vf_final needs mu);
    // 0 keeps them in vf_prep on the caller's stream (no auxiliary streams).
    int fork_mu;
};
// fork_mu of the verification entry points for a batch of n items: on for
// card-filling batches (65,536 Dilithium5 items 4.85 -> 4.78 ms, 10,000
// Dilithium2 items 1.33 -> 1.26 ms per call); off below 8,192 items, where
// the extra streams cost concurrent callers more than the overlap gives
// (serving saturation -30 %, ten 1,000-item callers -10 %).  The entry point
// that allocates nothing keeps 0: it creates no auxiliary streams.
#ifndef DIL_VERIFY_FORK_MU
#define DIL_VERIFY_FORK_MU 1
#endif
constexpr size_t VERIFY_FORK_MIN_ITEMS = 8192;
inline int verify_fork_mu(size_t n) { return DIL_VERIFY_FORK_MU && n >= VERIFY_FORK_MIN_ITEMS ? 1 : 0; }

// Host-side per-item state of the per-stage scheduler, laid out as the
// reference's sign_lut_element (api.cuh:31-34).
struct SignLutEntry {
    int done_flag;
    uint16_t nonce;
};

struct SignArgs;
struct SlotRef;

int launch_keygen(int mode, const KeygenArgs &a, hipStream_t st);
// Expands keys [0, n) of a key store (KeyLayout rows from `keys`): tr, A and
// t1^ from the public key, and s1^, s2^, t0^ when with_sk.
int launch_key_expand(int mode, uint8_t *keys, size_t key_pitch, size_t n, int with_sk, hipStream_t st);
// ev: null, or four events recorded around vf_prep / vf_main / vf_final
int launch_verify(int mode, const VerifyArgs &a, hipStream_t st, hipEvent_t *ev = nullptr);
int launch_sign_fixed(int mode, const SignArgs &a, hipStream_t st);
// mu / rhoprime (signing) and tr / mu (verification) with one sponge per
// 16-lane row instead of one lane: for batches of long messages
int launch_sign_mu_row(int mode, const SignArgs &a, hipStream_t st);
int launch_verify_mu_row(int mode, const VerifyArgs &a, hipStream_t st);
int launch_sign_round(int mode, const SignArgs &a, const SlotRef &sr, const uint32_t *d_exec, size_t nslot,
                      int by_item, uint8_t *d_done, hipStream_t st);
int launch_sign_copy(int mode, const SignArgs &a, const SlotRef &sr, const uint64_t *d_copy, size_t ncopy,
                     hipStream_t st);
// Status bits of a signing call (the control word after the queue head).
constexpr uint32_t SIGN_STATUS_ROUND_CAP = 1u;  // a workgroup stopped at its round limit
constexpr uint32_t SIGN_STATUS_NO_SECRET = 2u;  // an item's resident key has no secret key
// Resident keys: zeroes the signatures of items whose key has no secret key
// and sets SIGN_STATUS_NO_SECRET in *status (no-op without key indices).
int launch_sign_key_check(int mode, const SignArgs &a, uint32_t *status, hipStream_t st);
// Clears the control words d_head[0..2] (a kernel, so graph capture records
// no memset node).
int launch_zero_ctl(uint32_t *d_head, hipStream_t st);
int sign_perstage(int mode, const SignArgs &a, const SlotRef &item_slots, const SlotRef &temp_slots, size_t thr,
                  SignLutEntry *sign_lut, uint32_t *h_exec, uint32_t *d_exec, uint8_t *h_done, uint8_t *d_done,
                  uint64_t *h_copy, uint64_t *d_copy, hipStream_t st);
// Rejection loop of the persistent engine over nslots slots (the fixed part
// must have been launched first).  blocks <= 0: nslots / slots_per_block.
int launch_sign_persistent(int mode, const SignArgs &a, const SlotRef &sr, size_t nslots, int slots_per_block,
                           int occupancy, int blocks, int claim, uint32_t *d_head, unsigned long long *phase_sums,
                           hipStream_t st);
int default_slots_per_block(size_t nslots);

}  // namespace dil
