// Verification that allocates nothing: dil_verify_device with the per-item
// work rows (mu || w1', challenge masks, hint bitmap, status; VerifyLayout
// [HASH, M)) in a fixed device array of this library instead of a caller
// workspace, for callers that cannot allocate device memory (code under
// HIP-graph capture, serving loops that forbid hipMalloc).
//
// The array is a file-scope __device__ symbol, one copy per GPU, held until
// the process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress on the first call there (with that device current) and
// kept; kernels receive that address, never the symbol's host-side name.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race.  Nothing in it needs clearing between calls: each kernel of the
// pipeline writes every field it or a later kernel reads for an item before
// reading it (vf_prep clears the hint bitmap and writes the status first).
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/cudilithium_amd.h"
#include "internal.hpp"
#include "layout.hpp"
#include "pipelines.hpp"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)DIL_VERIFY_NOALLOC_SCRATCH_BYTES;

__device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
uint8_t *g_scratch_dev[64] = {nullptr};

// device address of the current device's copy of `scratch`
int scratch_address(uint8_t **out) {
    int dev = 0;
    int e = (int)hipGetDevice(&dev);
    if (e) return e;
    if (dev < 0 || dev >= 64) return DIL_ERANGE;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        if ((e = (int)hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_scratch_dev[dev] = static_cast<uint8_t *>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

size_t work_bytes(int mode) {
    switch (mode) {
        case 2: return dil::VerifyLayout<dil::Params<2>>::WORK_BYTES;
        case 3: return dil::VerifyLayout<dil::Params<3>>::WORK_BYTES;
        case 5: return dil::VerifyLayout<dil::Params<5>>::WORK_BYTES;
        default: return 0;
    }
}

size_t work_begin(int mode) {
    switch (mode) {
        case 2: return dil::VerifyLayout<dil::Params<2>>::WORK_BEGIN;
        case 3: return dil::VerifyLayout<dil::Params<3>>::WORK_BEGIN;
        default: return dil::VerifyLayout<dil::Params<5>>::WORK_BEGIN;
    }
}

}  // namespace

extern "C" {

size_t dil_verify_noalloc_max_items(int mode) {
    const size_t w = work_bytes(mode);
    return w ? NOALLOC_SCRATCH_BYTES / w : 0;
}

int dil_verify_device_noalloc(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_pk,
                              size_t pk_pitch, const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                              size_t n, void *stream) {
    if (mode != 2 && mode != 3 && mode != 5) return DIL_EINVAL;
    if (n == 0) return 0;
    if (!d_ret || !d_sig || !d_pk || !d_moff || !d_mlen || !d_msg) return DIL_EINVAL;
    if (sig_pitch < dil_signature_bytes(mode) || pk_pitch < dil_public_key_bytes(mode)) return DIL_EINVAL;
    if (n > dil_verify_noalloc_max_items(mode)) return DIL_ERANGE;
    uint8_t *base = nullptr;
    const int e = scratch_address(&base);
    if (e) return e;
    // rows of WORK_BYTES whose [HASH, M) part is item i's slice of the array
    // (the kernels never touch a row below HASH)
    const size_t pitch = work_bytes(mode);
    dil::VerifyArgs a{d_sig, sig_pitch, d_pk, pk_pitch, dil::MsgRef{d_msg, d_moff, 0, d_mlen, 0},
                      base - work_begin(mode), pitch, d_ret, 1, n};
    return dil::launch_verify(mode, a, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
