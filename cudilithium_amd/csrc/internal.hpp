// Host-side helpers shared by the kernel translation units and the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "keccak.hpp"
#include "params.hpp"

namespace dil {

// Number of compute units of the current device (cached per device).
int device_cu_count();

// Grid size for a grid-stride kernel: enough blocks to cover `units` work
// items at `per_block` items per block, capped at `waves_per_cu` resident
// blocks per CU.
int default_grid_blocks(size_t units, int per_block, int blocks_per_cu = 8);

// Two auxiliary streams (default priority, non-blocking) and fork/join
// events of the current device for work forked from the caller's stream
// `caller`, created on first use and kept until the process ends.  Each
// caller stream gets a set of its own (up to a cap, then sets are shared),
// so forks from different streams do not queue behind each other.  A caller
// holds `lock` while it enqueues a fork/join so that host threads sharing a
// set cannot interleave their event records.
struct AuxStreams {
    hipStream_t s[2];
    hipEvent_t fork, join[2];
    std::mutex lock;
};
AuxStreams *aux_streams(hipStream_t caller);

// Kernel tilings chosen at run time (set by the autotuner's cache through
// dil_set_kernel_tiling): per (kind, mode, batch bucket) a value, 0 for the
// compiled default.  Buckets are batch sizes up to 1K, 4K, 16K and above.
enum TilingKind { TILE_KEYGEN_ITEMS = 0, TILE_VERIFY_ITEMS = 1, TILE_NTT_WAVES = 2, TILE_KINDS = 3 };
constexpr int TILE_BUCKETS = 4;
int tiling_bucket(size_t n);
int tiling_get(int kind, int mode, size_t n);  // 0: the compiled default
int tiling_set(int kind, int mode, int bucket, int value);
// whether a value (0 = default) is one the kernels are built for
bool keygen_items_supported(int mode, int g);
bool verify_items_supported(int mode, int g);
bool ntt_waves_supported(int w);

// Launchers (return a hipError_t value, 0 on success).
int launch_ntt(int kind, int32_t *c, const int32_t *a, const int32_t *b, size_t npoly, size_t pitch_words,
               int blocks, hipStream_t stream);
int launch_keccak(uint64_t *st, size_t n, hipStream_t stream);
// a stream with a hardware queue of its own (all-CU mask)
int create_queue_stream(hipStream_t *st);
int valu_microbench(int op, uint32_t iters, double *ops_per_s);
int keccak_microbench(int form, uint32_t blocks, uint32_t threads, uint32_t nperm, double *us_per_perm,
                      double *perms_per_s);

}  // namespace dil
