// C-ABI layer of libcudilithium_amd.so: argument checking, pool layouts and
// the host side of each entry point declared in include/cudilithium_amd.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <atomic>
#include <functional>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/cudilithium_amd.h"
#include "internal.hpp"

namespace dil {

int device_cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached[dev] = cus;
    }
    return cached[dev];
}

// A stream with a hardware queue of its own.  The HIP runtime may map several
// streams created while the device is idle onto one hardware queue, where
// their work serialises (a short batch then waits behind a long one); a
// stream created with a CU mask always gets its own queue, and the mask here
// enables every CU, so nothing else changes.
int create_queue_stream(hipStream_t *st) {
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return (int)hipStreamCreateWithFlags(st, hipStreamNonBlocking);
    std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0xFFFFFFFFu);
    if (cus % 32) mask.back() = (1u << (cus % 32)) - 1u;
    return (int)hipExtStreamCreateWithCUMask(st, (uint32_t)mask.size(), mask.data());
}

AuxStreams *aux_streams(hipStream_t caller) {
    // per device: one set per caller stream (up to kAuxSets, then shared by
    // hash), so that callers on different streams fork onto different
    // auxiliary streams and their fixed stages run concurrently on the GPU
    constexpr size_t kAuxSets = 16;
    struct PerDevice {
        std::vector<std::pair<hipStream_t, AuxStreams *>> sets;
    };
    static PerDevice per_dev[64];
    static std::mutex mu;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    auto &v = per_dev[dev].sets;
    for (auto &e : v)
        if (e.first == caller) return e.second;
    if (v.size() >= kAuxSets) return v[std::hash<void *>()(reinterpret_cast<void *>(caller)) % v.size()].second;
    AuxStreams *a = new AuxStreams();
    // plain streams: with a hardware queue per auxiliary stream as well, ten
    // concurrent callers ran their signing calls 4x slower (too many queues)
    bool ok = hipStreamCreateWithFlags(&a->s[0], hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&a->s[1], hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&a->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&a->join[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&a->join[1], hipEventDisableTiming) == hipSuccess;
    if (!ok) return nullptr;  // (a partly created set is not reused)
    v.emplace_back(caller, a);
    return a;
}

namespace {
// [kind][mode slot 0..2 for modes 2/3/5][bucket]
std::atomic<int> g_tiling[TILE_KINDS][3][TILE_BUCKETS];

int mode_slot(int mode) { return mode == 2 ? 0 : mode == 3 ? 1 : mode == 5 ? 2 : -1; }
}  // namespace

int tiling_bucket(size_t n) { return n <= 1024 ? 0 : n <= 4096 ? 1 : n <= 16384 ? 2 : 3; }

int tiling_get(int kind, int mode, size_t n) {
    const int ms = kind == TILE_NTT_WAVES ? 0 : mode_slot(mode);
    if (kind < 0 || kind >= TILE_KINDS || ms < 0) return 0;
    return g_tiling[kind][ms][tiling_bucket(n)].load(std::memory_order_relaxed);
}

int tiling_set(int kind, int mode, int bucket, int value) {
    const int ms = kind == TILE_NTT_WAVES ? 0 : mode_slot(mode);
    if (kind < 0 || kind >= TILE_KINDS || ms < 0 || bucket < -1 || bucket >= TILE_BUCKETS || value < 0) return -1;
    for (int b = 0; b < TILE_BUCKETS; ++b)
        if (bucket < 0 || b == bucket) g_tiling[kind][ms][b].store(value, std::memory_order_relaxed);
    return 0;
}

int default_grid_blocks(size_t units, int per_block, int blocks_per_cu) {
    size_t need = (units + per_block - 1) / per_block;
    size_t cap = (size_t)device_cu_count() * blocks_per_cu;
    if (need > cap) need = cap;
    if (need == 0) need = 1;
    return (int)need;
}

}  // namespace dil

using namespace dil;

static inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

int dil_set_kernel_tiling(int kind, int mode, int bucket, int value) {
    bool ok = false;
    switch (kind) {
        case DIL_TILE_KEYGEN_ITEMS: ok = keygen_items_supported(mode, value); break;
        case DIL_TILE_VERIFY_ITEMS: ok = verify_items_supported(mode, value); break;
        case DIL_TILE_NTT_WAVES: ok = ntt_waves_supported(value); break;
        default: ok = false;
    }
    if (!ok) return DIL_EINVAL;
    return tiling_set(kind, mode, bucket, value) ? DIL_EINVAL : 0;
}

int dil_get_kernel_tiling(int kind, int mode, size_t n) { return tiling_get(kind, mode, n); }

size_t dil_public_key_bytes(int mode) {
    switch (mode) {
        case 2: return Sizes<Params<2>>::PK_BYTES;
        case 3: return Sizes<Params<3>>::PK_BYTES;
        case 5: return Sizes<Params<5>>::PK_BYTES;
        default: return 0;
    }
}

size_t dil_secret_key_bytes(int mode) {
    switch (mode) {
        case 2: return Sizes<Params<2>>::SK_BYTES;
        case 3: return Sizes<Params<3>>::SK_BYTES;
        case 5: return Sizes<Params<5>>::SK_BYTES;
        default: return 0;
    }
}

size_t dil_signature_bytes(int mode) {
    switch (mode) {
        case 2: return Sizes<Params<2>>::SIG_BYTES;
        case 3: return Sizes<Params<3>>::SIG_BYTES;
        case 5: return Sizes<Params<5>>::SIG_BYTES;
        default: return 0;
    }
}

/* ---------------------------- NTT --------------------------------------- */

int dil_ntt_forward(int32_t *d_poly, size_t npoly, size_t pitch_words, void *stream) {
    if ((npoly && !d_poly) || pitch_words < (size_t)N || (pitch_words & 3)) return DIL_EINVAL;
    return launch_ntt(0, d_poly, d_poly, nullptr, npoly, pitch_words, 0, S(stream));
}

int dil_ntt_inverse(int32_t *d_poly, size_t npoly, size_t pitch_words, void *stream) {
    if ((npoly && !d_poly) || pitch_words < (size_t)N || (pitch_words & 3)) return DIL_EINVAL;
    return launch_ntt(1, d_poly, d_poly, nullptr, npoly, pitch_words, 0, S(stream));
}

int dil_ntt_pointwise_pipeline(int32_t *d_c, const int32_t *d_a, const int32_t *d_b, size_t npoly,
                               size_t pitch_words, void *stream) {
    if ((npoly && (!d_c || !d_a || !d_b)) || pitch_words < (size_t)N || (pitch_words & 3)) return DIL_EINVAL;
    return launch_ntt(2, d_c, d_a, d_b, npoly, pitch_words, 0, S(stream));
}

int dil_valu_microbench(int op, uint32_t iters, double *lane_ops_per_s) {
    if (op < 0 || op > 7 || !lane_ops_per_s || iters == 0) return DIL_EINVAL;
    return valu_microbench(op, iters, lane_ops_per_s);
}

int dil_keccak_microbench(int form, uint32_t blocks, uint32_t threads, uint32_t nperm, double *us_per_perm,
                          double *perms_per_s) {
    if (!us_per_perm || !perms_per_s || blocks > (1u << 20)) return DIL_EINVAL;
    const int e = keccak_microbench(form, blocks, threads, nperm, us_per_perm, perms_per_s);
    return e == -1 ? DIL_EINVAL : e;
}

int dil_keccak_f1600_batch(uint64_t *d_states, size_t n, void *stream) {
    if (n && !d_states) return DIL_EINVAL;
    return launch_keccak(d_states, n, S(stream));
}

/* ---------------------------- runtime ----------------------------------- */

int dil_device_count(int *count) { return (int)hipGetDeviceCount(count); }
int dil_set_device(int device) { return (int)hipSetDevice(device); }

int dil_device_name(char *buf, size_t len) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return (int)e;
    if (buf && len) {
        snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    }
    return 0;
}

int dil_device_cu_count(void) { return device_cu_count(); }
int dil_malloc_device(void **ptr, size_t bytes) { return (int)hipMalloc(ptr, bytes ? bytes : 1); }
int dil_malloc_pitch(void **ptr, size_t *pitch, size_t width, size_t height) {
    return (int)hipMallocPitch(ptr, pitch, width ? width : 1, height ? height : 1);
}
int dil_free_device(void *ptr) { return (int)hipFree(ptr); }
int dil_malloc_host(void **ptr, size_t bytes) { return (int)hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault); }
int dil_free_host(void *ptr) { return (int)hipHostFree(ptr); }
int dil_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    return bytes ? (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)) : 0;
}
int dil_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    return bytes ? (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)) : 0;
}
int dil_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    return bytes ? (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)) : 0;
}
int dil_memset_device(void *dst, int value, size_t bytes, void *stream) {
    return bytes ? (int)hipMemsetAsync(dst, value, bytes, S(stream)) : 0;
}
// Streams for callers get a hardware queue each: the runtime's default is a
// few queues shared by all streams, where concurrent callers' kernels
// serialise (ten host threads signing 1,000-item batches on ten streams:
// 10.4 -> 5.0 ms per round of calls).
#ifndef DIL_STREAM_OWN_QUEUE
#define DIL_STREAM_OWN_QUEUE 1
#endif
int dil_stream_create(void **stream) {
    hipStream_t s;
    const int e = DIL_STREAM_OWN_QUEUE ? create_queue_stream(&s) : (int)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == 0) *stream = s;
    return e;
}
int dil_stream_destroy(void *stream) { return (int)hipStreamDestroy(S(stream)); }
int dil_stream_synchronize(void *stream) { return (int)hipStreamSynchronize(S(stream)); }
int dil_device_synchronize(void) { return (int)hipDeviceSynchronize(); }
int dil_event_create(void **event) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r == hipSuccess) *event = e;
    return (int)r;
}
int dil_event_destroy(void *event) { return (int)hipEventDestroy(reinterpret_cast<hipEvent_t>(event)); }
int dil_event_record(void *event, void *stream) {
    return (int)hipEventRecord(reinterpret_cast<hipEvent_t>(event), S(stream));
}
int dil_event_synchronize(void *event) { return (int)hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)); }
int dil_event_elapsed_ms(float *ms, void *start, void *stop) {
    return (int)hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop));
}

/* ---------------------------- synthetic inputs --------------------------- */

// xoshiro256** (Blackman-Vigna) seeded through splitmix64; the workload
// generator of the benchmark configurations.
int dil_xoshiro_bytes(uint64_t seed, uint8_t *out, size_t n) {
    if (n && !out) return DIL_EINVAL;
    uint64_t x = seed, s[4];
    for (int i = 0; i < 4; ++i) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        s[i] = z ^ (z >> 31);
    }
    auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
    for (size_t pos = 0; pos < n; pos += 8) {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        const size_t m = n - pos < 8 ? n - pos : 8;
        for (size_t b = 0; b < m; ++b) out[pos + b] = (uint8_t)(r >> (8 * b));
    }
    return 0;
}

const char *dil_error_string(int code) {
    switch (code) {
        case DIL_OK: return "success";
        case DIL_EINVAL: return "invalid argument";
        case DIL_ENOMEM: return "caller-provided buffer too small";
        case DIL_ERANGE: return "size beyond what the engine supports";
        default: return code > 0 ? hipGetErrorString((hipError_t)code) : "unknown error";
    }
}

}  // extern "C"
