// C-ABI entry points of the keygen / sign / verify pipelines: the drop-in
// replacements of the reference's api.cuh surface and the device-resident
// extensions.
#include <hip/hip_runtime.h>
#include <sys/random.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/cudilithium_amd.h"
#include "internal.hpp"
#include "layout.hpp"
#include "pipelines.hpp"
#include "sign_stages.hpp"

using namespace dil;

namespace {

hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

bool valid_mode(int m) { return m == 2 || m == 3 || m == 5; }

template <class F>
auto with_params(int mode, F &&f) {
    switch (mode) {
        case 2: return f(Params<2>{});
        case 3: return f(Params<3>{});
        default: return f(Params<5>{});
    }
}

// Pitched row copy done by the GPU, G bytes per access; one side may be
// pinned host memory reached through its device mapping.
template <int G>
__global__ void __launch_bounds__(256) rows_copy_kernel(uint8_t *dst, size_t dpitch, const uint8_t *src,
                                                        size_t spitch, size_t wg, size_t total) {
    using T = typename std::conditional<G == 16, uint4, typename std::conditional<G == 4, uint32_t, uint8_t>::type>::type;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t r = i / wg, c = (i - r * wg) * G;
        *reinterpret_cast<T *>(dst + r * dpitch + c) = *reinterpret_cast<const T *>(src + r * spitch + c);
    }
}

// Where a host-side pointer of a copy lives: pageable memory (unknown to the
// runtime), pinned host memory (hipHostMalloc) whose allocation holds
// [p, p + len) -- *view is then its device address -- or anything else.
enum class HostMem { pageable, pinned, other };

HostMem host_mem(const void *p, size_t len, const uint8_t **view) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is reported as an error; it is not one here
        return HostMem::pageable;
    }
    if (a.type == hipMemoryTypeUnregistered) return HostMem::pageable;
    if (a.type != hipMemoryTypeHost || !a.devicePointer) return HostMem::other;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, a.devicePointer) != hipSuccess) {
        (void)hipGetLastError();
        return HostMem::other;
    }
    const auto *d = static_cast<const uint8_t *>(a.devicePointer), *b = static_cast<const uint8_t *>(base);
    if (d < b || (size_t)(d - b) > size || len > size - (size_t)(d - b)) return HostMem::other;
    *view = d;
    return HostMem::pinned;
}

// d/s are device addresses (device memory or the device view of pinned host
// memory); width <= both pitches.
int rows_copy(uint8_t *d, size_t dpitch, const uint8_t *s, size_t spitch, size_t width, size_t height,
              hipStream_t st) {
    const uintptr_t all = reinterpret_cast<uintptr_t>(s) | reinterpret_cast<uintptr_t>(d) | spitch | dpitch | width;
    const int g = (all & 15) == 0 ? 16 : (all & 3) == 0 ? 4 : 1;
    const size_t wg = width / g, total = wg * height;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 2048);
    if (g == 16)
        rows_copy_kernel<16><<<blocks, 256, 0, st>>>(d, dpitch, s, spitch, wg, total);
    else if (g == 4)
        rows_copy_kernel<4><<<blocks, 256, 0, st>>>(d, dpitch, s, spitch, wg, total);
    else
        rows_copy_kernel<1><<<blocks, 256, 0, st>>>(d, dpitch, s, spitch, wg, total);
    return (int)hipGetLastError();
}

// Pinned staging for pageable host rows, which the GPU cannot reach: the rows
// are moved in chunks by a CPU memcpy plus the row-copy kernel, alternating
// between the buffer's two halves so that the CPU fills (or drains) one while
// the GPU drains (or fills) the other. One buffer per process (fine-grained,
// so the GPU never caches a stale chunk), allocated on first use and held
// until the process ends; one copy uses it at a time.
constexpr size_t STAGE_BYTES = 32ull << 20, STAGE_HALF = STAGE_BYTES / 2;
std::mutex g_stage_mu;
uint8_t *g_stage = nullptr, *g_stage_dev = nullptr;

int stage_ready() {
    if (g_stage) return 0;
    void *p = nullptr, *d = nullptr;
    int e = (int)hipHostMalloc(&p, STAGE_BYTES, hipHostMallocMapped | hipHostMallocCoherent);
    if (e) return e;
    if ((e = (int)hipHostGetDevicePointer(&d, p, 0))) {
        (void)hipHostFree(p);
        return e;
    }
    g_stage = static_cast<uint8_t *>(p);
    g_stage_dev = static_cast<uint8_t *>(d);
    return 0;
}

// nr rows of width bytes from `from` (pitch fp) to `to` (pitch tp); chunks of
// 8 MiB and more are split over PACK_THREADS host threads (one core's memcpy
// runs well below what the host memory and the GPU copy sustain)
constexpr unsigned PACK_THREADS = 4;

void pack_rows_serial(uint8_t *to, size_t tp, const uint8_t *from, size_t fp, size_t width, size_t nr) {
    if (tp == width && fp == width)
        std::memcpy(to, from, nr * width);
    else
        for (size_t i = 0; i < nr; ++i) std::memcpy(to + i * tp, from + i * fp, width);
}

void pack_rows(uint8_t *to, size_t tp, const uint8_t *from, size_t fp, size_t width, size_t nr) {
    if (nr * width < (8u << 20) || nr < PACK_THREADS) return pack_rows_serial(to, tp, from, fp, width, nr);
    auto part = [&](unsigned t) {
        const size_t r0 = nr * t / PACK_THREADS, r1 = nr * (t + 1) / PACK_THREADS;
        pack_rows_serial(to + r0 * tp, tp, from + r0 * fp, fp, width, r1 - r0);
    };
    std::thread th[PACK_THREADS - 1];
    for (unsigned t = 1; t < PACK_THREADS; ++t) {
        try {
            th[t - 1] = std::thread(part, t);
        } catch (...) {
            part(t);  // no thread available: copy this part here
        }
    }
    part(0);
    for (auto &x : th)
        if (x.joinable()) x.join();
}

int staged_copy(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height, bool h2d,
                hipStream_t st) {
    std::lock_guard<std::mutex> lock(g_stage_mu);
    int e = stage_ready();
    if (e) return e;
    // on an early return a copy kernel may still be queued on the stream with
    // a staging half as its source or destination: drain the stream before
    // the lock is released (its result is the error already being returned)
    struct Drain {
        hipStream_t st;
        bool armed = true;
        ~Drain() {
            if (armed) (void)hipStreamSynchronize(st);
        }
    } drain{st};
    const size_t rows_per = STAGE_HALF / width, nchunks = (height + rows_per - 1) / rows_per;
    auto rows_of = [&](size_t k) { return std::min(rows_per, height - k * rows_per); };
    auto host_half = [&](size_t k) { return g_stage + (k & 1) * STAGE_HALF; };
    auto dev_half = [&](size_t k) { return g_stage_dev + (k & 1) * STAGE_HALF; };
    if (h2d) {
        const auto *hs = static_cast<const uint8_t *>(src);
        auto *dd = static_cast<uint8_t *>(dst);
        for (size_t k = 0; k < nchunks; ++k) {
            // half k & 1 last held chunk k - 2, whose copy the previous
            // iteration's synchronisation has seen finish
            pack_rows(host_half(k), width, hs + k * rows_per * spitch, spitch, width, rows_of(k));
            if (k && (e = (int)hipStreamSynchronize(st))) return e;
            if ((e = rows_copy(dd + k * rows_per * dpitch, dpitch, dev_half(k), width, width, rows_of(k), st)))
                return e;
        }
        drain.armed = false;
        return (int)hipStreamSynchronize(st);  // the next user may overwrite the buffer
    }
    const auto *ds = static_cast<const uint8_t *>(src);
    auto *hd = static_cast<uint8_t *>(dst);
    for (size_t k = 0; k <= nchunks; ++k) {
        // chunk k - 1 is complete once the stream has drained; then the GPU
        // fills the other half with chunk k while the CPU unpacks chunk k - 1
        if (k && (e = (int)hipStreamSynchronize(st))) return e;
        if (k < nchunks &&
            (e = rows_copy(dev_half(k), width, ds + k * rows_per * spitch, spitch, width, rows_of(k), st)))
            return e;
        if (k) pack_rows(hd + (k - 1) * rows_per * dpitch, dpitch, host_half(k - 1), width, width, rows_of(k - 1));
    }
    drain.armed = false;  // the last iteration synchronised after the last copy
    return 0;
}

// Host <-> pitched device rows. hipMemcpy2DAsync moves a host-side 2D copy one
// row at a time (about 2 us per row: ~100 ms for 16,384 Dil3 signing rows from
// pinned memory, ~400 ms from pageable), so the GPU copies the rows itself:
// directly from / to pinned host memory (asynchronous, like the runtime copy),
// through the staging buffer for pageable memory (returns once copied, as the
// runtime's pageable copies do).
int copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height, hipMemcpyKind k,
           hipStream_t st) {
    if (width == 0 || height == 0) return 0;
    if ((k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToHost) && width <= spitch && width <= dpitch) {
        const bool h2d = k == hipMemcpyHostToDevice;
        const size_t hpitch = h2d ? spitch : dpitch;
        const uint8_t *hv = nullptr;
        const HostMem hm = host_mem(h2d ? src : dst, (height - 1) * hpitch + width, &hv);
        if (hm == HostMem::pinned) {
            if (h2d) return rows_copy(static_cast<uint8_t *>(dst), dpitch, hv, spitch, width, height, st);
            return rows_copy(const_cast<uint8_t *>(hv), dpitch, static_cast<const uint8_t *>(src), spitch, width,
                             height, st);
        }
        if (hm == HostMem::pageable && width <= STAGE_HALF)
            return staged_copy(dst, dpitch, src, spitch, width, height, h2d, st);
    }
    return (int)hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, k, st);
}

// Kind of a copy between the library's device rows and a caller pointer that
// may be host or device memory (the key store's load and store).
hipMemcpyKind caller_kind(const void *caller, size_t len, hipMemcpyKind host_kind) {
    const uint8_t *v = nullptr;
    return host_mem(caller, len, &v) == HostMem::other ? hipMemcpyDefault : host_kind;
}

}  // namespace

static_assert(sizeof(dil_sign_lut_element) == sizeof(SignLutEntry), "sign_lut_element layout");
static_assert(sizeof(dil_copy_lut_element) == 2 * sizeof(uint64_t), "copy_lut_element layout");

extern "C" {

/* ------------------------------------------------------- staging buffer */

int dil_release_host_staging(void) {
    // waits for a copy that holds the buffer; the next pageable copy
    // allocates it again
    std::lock_guard<std::mutex> lock(g_stage_mu);
    if (!g_stage) return 0;
    const int e = (int)hipHostFree(g_stage);
    g_stage = nullptr;
    g_stage_dev = nullptr;
    return e;
}

/* ------------------------------------------------------------ pool sizes */

size_t dil_keypair_pool_row_bytes(int mode) {
    if (!valid_mode(mode)) return 0;
    return with_params(mode, [](auto p) { return KeypairLayout<decltype(p)>::ROW; });
}

size_t dil_sign_pool_row_bytes(int mode, size_t mlen) {
    if (!valid_mode(mode)) return 0;
    return with_params(mode, [&](auto p) { return SignLayout<decltype(p)>::row(mlen); });
}

size_t dil_sign_temp_row_bytes(int mode) {
    if (!valid_mode(mode)) return 0;
    return with_params(mode, [](auto p) { return TempLayout<decltype(p)>::ROW; });
}

size_t dil_verify_pool_row_bytes(int mode, size_t mlen) {
    if (!valid_mode(mode)) return 0;
    return with_params(mode, [&](auto p) { return VerifyLayout<decltype(p)>::row(mlen); });
}

size_t dil_reference_pool_row_bytes(int mode, int which, size_t mlen) {
    if (!valid_mode(mode)) return 0;
    return with_params(mode, [&](auto p) -> size_t {
        using P = decltype(p);
        switch (which) {
            case DIL_POOL_KEYPAIR: return KeypairLayout<P>::REF_ROW;
            case DIL_POOL_SIGN: return SignLayout<P>::ref_row(mlen);
            case DIL_POOL_SIGN_TEMP: return TempLayout<P>::REF_ROW;
            case DIL_POOL_VERIFY: return VerifyLayout<P>::ref_row(mlen);
            default: return 0;
        }
    });
}

/* ------------------------------------------------------- drop-in: keygen */

int dil_crypto_sign_keypair(int mode, uint8_t *pk, uint8_t *sk, uint8_t *d_pool, size_t pitch, size_t batch,
                            void *stream, size_t rand_index) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (batch == 0) return 0;
    if (!pk || !sk || !d_pool) return DIL_EINVAL;
    if (pitch < dil_keypair_pool_row_bytes(mode)) return DIL_ENOMEM;
    return with_params(mode, [&](auto p) -> int {
        using P = decltype(p);
        using KL = KeypairLayout<P>;
        using Z = Sizes<P>;
        const hipStream_t st = S(stream);
        KeygenArgs a{d_pool + KL::PK, pitch, d_pool + KL::SK, pitch, d_pool + KL::SCRATCH, pitch, nullptr, 0,
                     (uint64_t)rand_index, batch};
        int e = launch_keygen(mode, a, st);
        if (e) return e;
        // api.cu:55-58: asynchronous D2H on the caller's stream
        if ((e = copy2d(pk, Z::PK_BYTES, d_pool + KL::PK, pitch, Z::PK_BYTES, batch, hipMemcpyDeviceToHost, st)))
            return e;
        return copy2d(sk, Z::SK_BYTES, d_pool + KL::SK, pitch, Z::SK_BYTES, batch, hipMemcpyDeviceToHost, st);
    });
}

/* --------------------------------------------------------- drop-in: sign */

// rnd (randomized signing): batch x 64 host bytes of rhoprime, or null
static int sign_dropin(int mode, int engine, uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                       size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool, size_t sign_pitch,
                       uint8_t *d_temp_pool, size_t temp_pitch, dil_task_lut *lut, size_t exec_threshold,
                       size_t batch, void *stream, const uint8_t *rnd = nullptr) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (siglen) *siglen = dil_signature_bytes(mode);
    if (batch == 0) return 0;
    if (!sig || !sk || !d_sign_pool || !lut || exec_threshold == 0 || (mlen && !m)) return DIL_EINVAL;
    if (batch > 65536) return DIL_ERANGE;  // (item << 16 | nonce) entries, api.cu:224
    if (sign_pitch < dil_sign_pool_row_bytes(mode, mlen)) return DIL_ENOMEM;
    if (!d_temp_pool || temp_pitch < dil_sign_temp_row_bytes(mode)) return DIL_ENOMEM;
    // the rows' int32 arrays are read and written 16 bytes per lane (hipMallocPitch pitches qualify)
    if (sign_pitch % 16 || temp_pitch % 16 || (reinterpret_cast<uintptr_t>(d_sign_pool) & 15) ||
        (reinterpret_cast<uintptr_t>(d_temp_pool) & 15))
        return DIL_EINVAL;
    if (!lut->sign_lut || lut->sign_lut_len < batch || !lut->h_exec_lut || !lut->d_exec_lut || !lut->h_done_lut ||
        !lut->d_done_lut || !lut->h_copy_lut || !lut->d_copy_lut)
        return DIL_EINVAL;
    return with_params(mode, [&](auto p) -> int {
        using P = decltype(p);
        using SLy = SignLayout<P>;
        using Z = Sizes<P>;
        const hipStream_t st = S(stream);
        int e;
        // api.cu:117-126
        if ((e = copy2d(d_sign_pool + SLy::M, sign_pitch, m, m_pitch, mlen, batch, hipMemcpyHostToDevice, st))) return e;
        if ((e = copy2d(d_sign_pool + SLy::SK, sign_pitch, sk, Z::SK_BYTES, Z::SK_BYTES, batch, hipMemcpyHostToDevice,
                        st)))
            return e;
        // randomized signing: rhoprime rows into each row's rhoprime field,
        // which the fixed stage then reads instead of deriving it (api.cu:143-148)
        if (rnd && (e = copy2d(d_sign_pool + SLy::RHOPRIME, sign_pitch, rnd, CRHBYTES, CRHBYTES, batch,
                               hipMemcpyHostToDevice, st)))
            return e;
        SignArgs a{};
        a.sig = d_sign_pool + SLy::SIG;
        a.sig_pitch = sign_pitch;
        a.sk = d_sign_pool + SLy::SK;
        a.sk_pitch = sign_pitch;
        a.m = MsgRef{d_sign_pool + SLy::M, nullptr, sign_pitch, nullptr, (uint32_t)mlen};
        a.rows = d_sign_pool;
        a.row_pitch = sign_pitch;
        a.off_ints = SLy::ints(mlen);
        a.n = batch;
        if (rnd) {
            a.rnd = d_sign_pool + SLy::RHOPRIME;
            a.rnd_pitch = sign_pitch;
        }
        const SlotRef item_slots{d_sign_pool + SLy::SLOT, sign_pitch, d_sign_pool + SLy::SIG, sign_pitch};
        const SlotRef temp_slots{d_temp_pool + TempLayout<P>::SLOT, temp_pitch, d_temp_pool + TempLayout<P>::SIG,
                                 temp_pitch};
        if (engine == DIL_ENGINE_PERSTAGE) {
            e = sign_perstage(mode, a, item_slots, temp_slots, exec_threshold,
                              reinterpret_cast<SignLutEntry *>(lut->sign_lut), lut->h_exec_lut, lut->d_exec_lut,
                              lut->h_done_lut, lut->d_done_lut, reinterpret_cast<uint64_t *>(lut->h_copy_lut),
                              reinterpret_cast<uint64_t *>(lut->d_copy_lut), st);
        } else {
            // Attempt slots. Up to exec_threshold items: the temp rows. Larger
            // batches: slot s keeps its state in sign-pool row s (its slot area,
            // otherwise unused here) and its candidate signature in the temp pool
            // read as one array of SIG_STRIDE entries, so the same pools give up
            // to `batch` slots (a card-filling batch at the reference's default
            // exec_threshold of 2048). Results do not depend on the slot count.
            // The first word of d_copy_lut is the work-queue head.
            if ((e = launch_sign_fixed(mode, a, st))) return e;
            constexpr size_t SIG_STRIDE = align256(Z::SIG_BYTES);
            SlotRef slots = temp_slots;
            size_t nslots = exec_threshold;
            if (batch > exec_threshold) {
                nslots = std::min(batch, exec_threshold * temp_pitch / SIG_STRIDE);
                slots = SlotRef{d_sign_pool + SLy::SLOT, sign_pitch, d_temp_pool, SIG_STRIDE};
            }
            uint32_t *ctl = reinterpret_cast<uint32_t *>(lut->d_copy_lut);
            e = launch_sign_persistent(mode, a, slots, nslots, default_slots_per_block(nslots), 0, 0, 0, ctl, nullptr,
                                       st);
            if (e == -3) return DIL_ERANGE;
            if (e) return e;
            // like the reference's call, return once the rejection loop is done
            uint32_t status = 0;
            if ((e = (int)hipMemcpyAsync(&status, ctl + 1, sizeof(status), hipMemcpyDeviceToHost, st))) return e;
            if ((e = (int)hipStreamSynchronize(st))) return e;
            if (status) return DIL_ERANGE;
        }
        if (e) return e;
        // api.cu:443-446
        return copy2d(sig, sig_pitch, d_sign_pool + SLy::SIG, sign_pitch, Z::SIG_BYTES, batch, hipMemcpyDeviceToHost,
                      st);
    });
}

int dil_crypto_sign_signature(int mode, uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                              size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool,
                              size_t sign_pitch, uint8_t *d_temp_pool, size_t temp_pitch, dil_task_lut *lut,
                              size_t exec_threshold, size_t batch, void *stream) {
    return sign_dropin(mode, DIL_ENGINE_PERSISTENT, sig, sig_pitch, siglen, m, m_pitch, mlen, sk, d_sign_pool,
                       sign_pitch, d_temp_pool, temp_pitch, lut, exec_threshold, batch, stream);
}

int dil_crypto_sign_signature_perstage(int mode, uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                                       size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool,
                                       size_t sign_pitch, uint8_t *d_temp_pool, size_t temp_pitch,
                                       dil_task_lut *lut, size_t exec_threshold, size_t batch, void *stream) {
    return sign_dropin(mode, DIL_ENGINE_PERSTAGE, sig, sig_pitch, siglen, m, m_pitch, mlen, sk, d_sign_pool,
                       sign_pitch, d_temp_pool, temp_pitch, lut, exec_threshold, batch, stream);
}

namespace {

int os_randombytes(uint8_t *out, size_t outlen) {
    while (outlen) {
        const ssize_t r = getrandom(out, outlen, 0);
        if (r < 0) {
            if (errno == EINTR) continue;
            return DIL_EINVAL;
        }
        out += r;
        outlen -= (size_t)r;
    }
    return 0;
}

std::atomic<dil_randombytes_fn> g_randombytes{nullptr};

}  // namespace

void dil_set_randombytes(dil_randombytes_fn fn) { g_randombytes.store(fn); }

int dil_randombytes(uint8_t *out, size_t outlen) {
    if (!out && outlen) return DIL_EINVAL;
    const dil_randombytes_fn fn = g_randombytes.load();
    return fn ? fn(out, outlen) : os_randombytes(out, outlen);
}

int dil_crypto_sign_signature_randomized(int mode, uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                                         size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool,
                                         size_t sign_pitch, uint8_t *d_temp_pool, size_t temp_pitch,
                                         dil_task_lut *lut, size_t exec_threshold, size_t batch, void *stream) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (batch == 0) {
        if (siglen) *siglen = dil_signature_bytes(mode);
        return 0;
    }
    if (batch > 65536) return DIL_ERANGE;
    std::vector<uint8_t> rnd;
    try {
        rnd.resize(batch * CRHBYTES);
    } catch (...) {
        return DIL_ENOMEM;
    }
    const int e = dil_randombytes(rnd.data(), rnd.size());
    if (e) return e;
    // the staged copy of the pageable rows completes before sign_dropin returns
    return sign_dropin(mode, DIL_ENGINE_PERSISTENT, sig, sig_pitch, siglen, m, m_pitch, mlen, sk, d_sign_pool,
                       sign_pitch, d_temp_pool, temp_pitch, lut, exec_threshold, batch, stream, rnd.data());
}

/* ------------------------------------------------------- drop-in: verify */

int dil_crypto_sign_verify(int mode, int *ret, const uint8_t *sig, size_t sig_pitch, size_t siglen, const uint8_t *m,
                           size_t m_pitch, size_t mlen, const uint8_t *pk, uint8_t *d_pool, size_t pitch,
                           size_t batch, void *stream) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (batch == 0) return 0;
    if (!ret) return DIL_EINVAL;
    if (siglen != dil_signature_bytes(mode)) {  // api.cu:461-464
        for (size_t i = 0; i < batch; ++i) ret[i] = -1;
        return 0;
    }
    if (!sig || !pk || !d_pool || (mlen && !m)) return DIL_EINVAL;
    if (pitch < dil_verify_pool_row_bytes(mode, mlen)) return DIL_ENOMEM;
    if (pitch % 4) return DIL_EINVAL;  // the verdict is an int32 in each row
    return with_params(mode, [&](auto p) -> int {
        using P = decltype(p);
        using VL = VerifyLayout<P>;
        using Z = Sizes<P>;
        const hipStream_t st = S(stream);
        int e;
        if ((e = copy2d(d_pool + VL::SIG, pitch, sig, sig_pitch, Z::SIG_BYTES, batch, hipMemcpyHostToDevice, st)))
            return e;
        if ((e = copy2d(d_pool + VL::M, pitch, m, m_pitch, mlen, batch, hipMemcpyHostToDevice, st))) return e;
        if ((e = copy2d(d_pool + VL::PK, pitch, pk, Z::PK_BYTES, Z::PK_BYTES, batch, hipMemcpyHostToDevice, st)))
            return e;
        int32_t *d_ret = reinterpret_cast<int32_t *>(d_pool + VL::FLAG + 4);
        VerifyArgs a{d_pool + VL::SIG, pitch, d_pool + VL::PK, pitch,
                     MsgRef{d_pool + VL::M, nullptr, pitch, nullptr, (uint32_t)mlen},
                     d_pool, pitch, d_ret, pitch / 4, batch};
        a.fork_mu = verify_fork_mu(batch);
        if ((e = launch_verify(mode, a, st))) return e;
        // api.cu:513
        return copy2d(ret, sizeof(int), d_ret, pitch, sizeof(int), batch, hipMemcpyDeviceToHost, st);
    });
}

/* ------------------------------------------------ per-mode drop-in names */

#define DIL_MODE_WRAPPERS(M)                                                                                        \
    int dil##M##_crypto_sign_keypair(uint8_t *pk, uint8_t *sk, uint8_t *d_pool, size_t pitch, size_t batch,       \
                                     void *stream, size_t rand_index) {                                            \
        return dil_crypto_sign_keypair(M, pk, sk, d_pool, pitch, batch, stream, rand_index);                      \
    }                                                                                                               \
    int dil##M##_crypto_sign_signature(uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,           \
                                       size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool,       \
                                       size_t sign_pitch, uint8_t *d_temp_pool, size_t temp_pitch,                 \
                                       dil_task_lut *lut, size_t exec_threshold, size_t batch, void *stream) {     \
        return dil_crypto_sign_signature(M, sig, sig_pitch, siglen, m, m_pitch, mlen, sk, d_sign_pool, sign_pitch, \
                                         d_temp_pool, temp_pitch, lut, exec_threshold, batch, stream);             \
    }                                                                                                               \
    int dil##M##_crypto_sign_verify(int *ret, const uint8_t *sig, size_t sig_pitch, size_t siglen,                \
                                    const uint8_t *m, size_t m_pitch, size_t mlen, const uint8_t *pk,              \
                                    uint8_t *d_pool, size_t pitch, size_t batch, void *stream) {                   \
        return dil_crypto_sign_verify(M, ret, sig, sig_pitch, siglen, m, m_pitch, mlen, pk, d_pool, pitch, batch,  \
                                      stream);                                                                      \
    }

DIL_MODE_WRAPPERS(2)
DIL_MODE_WRAPPERS(3)
DIL_MODE_WRAPPERS(5)

/* ------------------------------------------------- device-resident batches */

size_t dil_keypair_work_row_bytes(int mode) {
    if (!valid_mode(mode)) return 0;
    return with_params(mode, [](auto p) {
        using KL = KeypairLayout<decltype(p)>;
        return KL::ROW - KL::SCRATCH;
    });
}

int dil_keypair_device(int mode, uint8_t *d_pk, size_t pk_pitch, uint8_t *d_sk, size_t sk_pitch,
                       const uint8_t *d_seeds, size_t seed_pitch, uint64_t rand_index, size_t n, uint8_t *d_work,
                       size_t work_pitch, void *stream) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (n == 0) return 0;
    if (!d_pk || !d_sk || !d_work) return DIL_EINVAL;
    if (pk_pitch < dil_public_key_bytes(mode) || sk_pitch < dil_secret_key_bytes(mode)) return DIL_EINVAL;
    if (work_pitch < dil_keypair_work_row_bytes(mode) || (work_pitch & 7)) return DIL_ENOMEM;
    if (d_seeds && seed_pitch < 32) return DIL_EINVAL;
    KeygenArgs a{d_pk, pk_pitch, d_sk, sk_pitch, d_work, work_pitch, d_seeds, seed_pitch, rand_index, n};
    return launch_keygen(mode, a, S(stream));
}

size_t dil_verify_work_row_bytes(int mode) { return dil_verify_pool_row_bytes(mode, 0); }

int dil_verify_device_timed(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_pk,
                            size_t pk_pitch, const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                            size_t n, uint8_t *d_work, size_t work_pitch, float *kernel_ms3, void *stream) {
    if (!kernel_ms3) return DIL_EINVAL;
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (n == 0) return 0;
    if (!d_ret || !d_sig || !d_pk || !d_work || !d_moff || !d_mlen || !d_msg) return DIL_EINVAL;
    if (sig_pitch < dil_signature_bytes(mode) || pk_pitch < dil_public_key_bytes(mode)) return DIL_EINVAL;
    if (work_pitch < dil_verify_work_row_bytes(mode) || (work_pitch & 255)) return DIL_ENOMEM;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int e = 0;
    for (auto &x : ev)
        if (!e) e = (int)hipEventCreate(&x);
    if (!e) {
        VerifyArgs a{d_sig, sig_pitch, d_pk, pk_pitch, MsgRef{d_msg, d_moff, 0, d_mlen, 0}, d_work, work_pitch, d_ret,
                     1, n};
        e = launch_verify(mode, a, S(stream), ev);
        if (!e) e = (int)hipEventSynchronize(ev[3]);
        for (int k = 0; k < 3 && !e; ++k) e = (int)hipEventElapsedTime(&kernel_ms3[k], ev[k], ev[k + 1]);
    }
    for (auto &x : ev)
        if (x) (void)hipEventDestroy(x);
    return e;
}

int dil_verify_device_ex(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_pk,
                         size_t pk_pitch, const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                         size_t n, uint8_t *d_work, size_t work_pitch, int flags, void *stream) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (n == 0) return 0;
    if (flags & ~DIL_HASH_LONG_MESSAGES) return DIL_EINVAL;
    if (!d_ret || !d_sig || !d_pk || !d_work || !d_moff || !d_mlen || !d_msg) return DIL_EINVAL;
    if (sig_pitch < dil_signature_bytes(mode) || pk_pitch < dil_public_key_bytes(mode)) return DIL_EINVAL;
    if (work_pitch < dil_verify_work_row_bytes(mode) || (work_pitch & 255)) return DIL_ENOMEM;
    VerifyArgs a{d_sig, sig_pitch, d_pk, pk_pitch, MsgRef{d_msg, d_moff, 0, d_mlen, 0}, d_work, work_pitch, d_ret, 1, n};
    a.row_mu = (flags & DIL_HASH_LONG_MESSAGES) ? 1 : 0;
    a.fork_mu = verify_fork_mu(n);
    return launch_verify(mode, a, S(stream));
}

int dil_verify_device(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_pk,
                      size_t pk_pitch, const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                      size_t n, uint8_t *d_work, size_t work_pitch, void *stream) {
    return dil_verify_device_ex(mode, d_ret, d_sig, sig_pitch, d_pk, pk_pitch, d_msg, d_moff, d_mlen, n, d_work,
                                work_pitch, 0, stream);
}

}  // extern "C"

/* ------------------------------------------------------ signing context */

struct dil_sign_ctx {
    int mode = 0;
    size_t max_batch = 0, nslots = 0;
    uint8_t *rows = nullptr;  // SignLayout rows (messages stay external)
    size_t row_pitch = 0;
    uint8_t *slots = nullptr;  // TempLayout rows = attempt slots
    size_t slot_pitch = 0;
    uint32_t *d_head = nullptr;
    // per-stage engine tables
    std::vector<SignLutEntry> sign_lut;
    uint32_t *h_exec = nullptr, *d_exec = nullptr;
    uint8_t *h_done = nullptr, *d_done = nullptr;
    uint64_t *h_copy = nullptr, *d_copy = nullptr;
    int slots_per_block = 0, blocks = 0, occupancy = 0;
    // timing: one (start, fixed done, loop done) event triple per timed call
    std::vector<hipEvent_t> ev;
    size_t timing_cap = 0, timed = 0;
    bool resident_only = false;  // rows without the per-item expanded arrays
    int claim = 0;               // items owned per workgroup (0 = slots per workgroup)
    bool profile = false;        // phase sums into d_head words 16..27
    bool row_mu = false;         // long messages: mu / rhoprime with one sponge per 16-lane row
};

// Resident keys: KeyLayout rows (pk, sk, tr, A, s1^, s2^, t0^, t1^).
struct dil_keystore {
    int mode = 0;
    size_t capacity = 0, pitch = 0;
    uint8_t *rows = nullptr;
    // host copy of each key's KEY_HAS_SECRET flag, set by load / keygen calls
    std::vector<uint8_t> has_sk;
};

// sign n items; with ks, item i uses resident key d_key_idx[i] (d_sk unused)
static int sign_enqueue(dil_sign_ctx *c, uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_sk, size_t sk_pitch,
                        const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, int engine,
                        hipStream_t st, bool timing, const dil_keystore *ks = nullptr,
                        const uint32_t *d_key_idx = nullptr, const uint8_t *d_rnd = nullptr,
                        size_t rnd_pitch = 0) {
    if (!c) return DIL_EINVAL;
    if (n == 0) return 0;
    if (n > c->max_batch) return DIL_ERANGE;
    if (ks) {
        if (ks->mode != c->mode || !d_key_idx || ks->capacity == 0) return DIL_EINVAL;
        d_sk = ks->rows + with_params(c->mode, [](auto p) { return KeyLayout<decltype(p)>::SK; });
        sk_pitch = ks->pitch;
    } else if (c->resident_only) {
        return DIL_EINVAL;
    }
    if (!d_sig || !d_sk || !d_msg || !d_moff || !d_mlen) return DIL_EINVAL;
    if (sig_pitch < dil_signature_bytes(c->mode) || sk_pitch < dil_secret_key_bytes(c->mode)) return DIL_EINVAL;
    if (engine != DIL_ENGINE_PERSISTENT && engine != DIL_ENGINE_PERSTAGE) return DIL_EINVAL;
    if (engine == DIL_ENGINE_PERSTAGE && n > 65536) return DIL_ERANGE;
    return with_params(c->mode, [&](auto p) -> int {
        using P = decltype(p);
        SignArgs a{};
        a.sig = d_sig;
        a.sig_pitch = sig_pitch;
        a.sk = d_sk;
        a.sk_pitch = sk_pitch;
        a.m = MsgRef{d_msg, d_moff, 0, d_mlen, 0};
        a.rows = c->rows;
        a.row_pitch = c->row_pitch;
        a.off_ints = SignLayout<P>::ints(0);
        a.n = n;
        a.rnd = d_rnd;
        a.rnd_pitch = rnd_pitch;
        a.row_mu = c->row_mu ? 1 : 0;
        if (ks) {
            a.key_idx = d_key_idx;
            a.key_ints = ks->rows + KeyLayout<P>::MAT;
            a.key_pitch = ks->pitch;
            a.nkeys = (uint32_t)ks->capacity;
        }
        const SlotRef slots{c->slots + TempLayout<P>::SLOT, c->slot_pitch, c->slots + TempLayout<P>::SIG,
                            c->slot_pitch};
        int e;
        hipEvent_t *e3 = timing && c->timed < c->timing_cap ? &c->ev[3 * c->timed] : nullptr;
        if (e3 && (e = (int)hipEventRecord(e3[0], st))) return e;
        if (engine == DIL_ENGINE_PERSTAGE) {
            // the per-stage scheduler launches the fixed part itself; the
            // status word is cleared here (the persistent launch clears it)
            if ((e = launch_zero_ctl(c->d_head, st))) return e;
            if (e3 && (e = (int)hipEventRecord(e3[1], st))) return e;
            const SlotRef items{c->rows + SignLayout<P>::SLOT, c->row_pitch, d_sig, sig_pitch};
            e = sign_perstage(c->mode, a, items, slots, c->nslots, c->sign_lut.data(), c->h_exec, c->d_exec,
                              c->h_done, c->d_done, c->h_copy, c->d_copy, st);
        } else {
            if ((e = launch_sign_fixed(c->mode, a, st))) return e;
            if (e3 && (e = (int)hipEventRecord(e3[1], st))) return e;
            const int spb = c->slots_per_block ? c->slots_per_block : default_slots_per_block(c->nslots);
            unsigned long long *ph =
                c->profile ? reinterpret_cast<unsigned long long *>(c->d_head + 16) : nullptr;
            e = launch_sign_persistent(c->mode, a, slots, c->nslots, spb, c->occupancy, c->blocks, c->claim, c->d_head,
                                       ph, st);
            if (e == -3) return DIL_ERANGE;
        }
        if (e) return e;
        if (ks && (e = launch_sign_key_check(c->mode, a, c->d_head + 1, st))) return e;
        if (e3) {
            if ((e = (int)hipEventRecord(e3[2], st))) return e;
            ++c->timed;
        }
        return 0;
    });
}


static int ctx_create(dil_sign_ctx **out, int mode, size_t max_batch, size_t nslots, bool resident_only) {
    if (!out || !valid_mode(mode) || max_batch == 0 || nslots == 0) return DIL_EINVAL;
    dil_sign_ctx *c = new dil_sign_ctx();
    c->mode = mode;
    c->max_batch = max_batch;
    c->nslots = nslots;
    c->resident_only = resident_only;
    c->row_pitch = resident_only
                       ? with_params(mode, [](auto p) { return align256(SignLayout<decltype(p)>::ints(0)); })
                       : align256(dil_sign_pool_row_bytes(mode, 0));
    c->slot_pitch = align256(dil_sign_temp_row_bytes(mode));
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    chk(hipMalloc((void **)&c->rows, c->row_pitch * max_batch));
    chk(hipMalloc((void **)&c->slots, c->slot_pitch * nslots));
    chk(hipMalloc((void **)&c->d_head, 256));
    chk(hipMalloc((void **)&c->d_exec, 4 * nslots));
    chk(hipMalloc((void **)&c->d_done, nslots));
    chk(hipMalloc((void **)&c->d_copy, 16 * nslots));
    chk(hipHostMalloc((void **)&c->h_exec, 4 * nslots, hipHostMallocDefault));
    chk(hipHostMalloc((void **)&c->h_done, nslots, hipHostMallocDefault));
    chk(hipHostMalloc((void **)&c->h_copy, 16 * nslots, hipHostMallocDefault));
    c->sign_lut.resize(max_batch);
    if (e != hipSuccess) {
        dil_sign_ctx_destroy(c);
        return (int)e;
    }
    *out = c;
    return 0;
}

extern "C" {

int dil_sign_ctx_destroy(dil_sign_ctx *c) {
    if (!c) return 0;
    (void)hipFree(c->rows);
    (void)hipFree(c->slots);
    (void)hipFree(c->d_head);
    (void)hipFree(c->d_exec);
    (void)hipFree(c->d_done);
    (void)hipFree(c->d_copy);
    (void)hipHostFree(c->h_exec);
    (void)hipHostFree(c->h_done);
    (void)hipHostFree(c->h_copy);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    delete c;
    return 0;
}

int dil_sign_ctx_create(dil_sign_ctx **out, int mode, size_t max_batch, size_t nslots) {
    return ctx_create(out, mode, max_batch, nslots, false);
}

int dil_sign_ctx_create_resident(dil_sign_ctx **out, int mode, size_t max_batch, size_t nslots) {
    return ctx_create(out, mode, max_batch, nslots, true);
}

int dil_sign_ctx_set_tiling(dil_sign_ctx *c, int slots_per_block, int occupancy, int blocks) {
    if (!c) return DIL_EINVAL;
    if (slots_per_block != 0 && slots_per_block != 1 && slots_per_block != 8 && slots_per_block != 16 &&
        slots_per_block != 32)
        return DIL_EINVAL;
    if (occupancy != 0 && (occupancy < 2 || occupancy > 4)) return DIL_EINVAL;
    if (slots_per_block > 0 && c->nslots / slots_per_block == 0) return DIL_ERANGE;
    c->slots_per_block = slots_per_block;
    c->occupancy = occupancy;
    c->blocks = blocks < 0 ? 0 : blocks;
    return 0;
}

int dil_sign_ctx_set_claim(dil_sign_ctx *c, int items_per_block) {
    if (!c || items_per_block < 0 || (items_per_block & ~(DIL_CLAIM_GUIDED | 0xFFFF))) return DIL_EINVAL;
    c->claim = items_per_block;
    return 0;
}

int dil_sign_ctx_enable_timing(dil_sign_ctx *c, int max_calls) {
    if (!c || max_calls < 0) return DIL_EINVAL;
    while (c->ev.size() < 3 * (size_t)max_calls) {
        hipEvent_t e;
        int r = (int)hipEventCreate(&e);
        if (r) return r;
        c->ev.push_back(e);
    }
    c->timing_cap = (size_t)max_calls;
    c->timed = 0;
    return 0;
}

int dil_sign_ctx_timed_calls(dil_sign_ctx *c) { return c ? (int)c->timed : 0; }

int dil_sign_ctx_timing(dil_sign_ctx *c, int call, float *fixed_ms, float *loop_ms) {
    if (!c || call < 0 || (size_t)call >= c->timed) return DIL_EINVAL;
    hipEvent_t *e3 = &c->ev[3 * (size_t)call];
    int e = (int)hipEventSynchronize(e3[2]);
    if (e) return e;
    if (fixed_ms && (e = (int)hipEventElapsedTime(fixed_ms, e3[0], e3[1]))) return e;
    if (loop_ms && (e = (int)hipEventElapsedTime(loop_ms, e3[1], e3[2]))) return e;
    return 0;
}

int dil_sign_ctx_status(dil_sign_ctx *c, int *status, void *stream) {
    if (!c || !status) return DIL_EINVAL;
    uint32_t v = 0;
    int e = (int)hipMemcpyAsync(&v, c->d_head + 1, sizeof(v), hipMemcpyDeviceToHost, S(stream));
    if (!e) e = (int)hipStreamSynchronize(S(stream));
    *status = (int)v;
    return e;
}

int dil_sign_ctx_executed_attempts(dil_sign_ctx *c, uint64_t *count, void *stream) {
    if (!c || !count) return DIL_EINVAL;
    uint32_t v = 0;
    int e = (int)hipMemcpyAsync(&v, c->d_head + 2, sizeof(v), hipMemcpyDeviceToHost, S(stream));
    if (!e) e = (int)hipStreamSynchronize(S(stream));
    *count = v;
    return e;
}

int dil_sign_ctx_set_message_hashing(dil_sign_ctx *c, int flags) {
    if (!c || (flags & ~DIL_HASH_LONG_MESSAGES)) return DIL_EINVAL;
    c->row_mu = (flags & DIL_HASH_LONG_MESSAGES) != 0;
    return 0;
}

int dil_sign_ctx_profile_phases(dil_sign_ctx *c, int enable) {
    if (!c) return DIL_EINVAL;
    c->profile = enable != 0;
    return 0;
}

int dil_sign_ctx_phase_cycles(dil_sign_ctx *c, uint64_t *cycles, void *stream) {
    if (!c || !cycles) return DIL_EINVAL;
    int e = (int)hipMemcpyAsync(cycles, c->d_head + 16, 6 * sizeof(uint64_t), hipMemcpyDeviceToHost, S(stream));
    return e ? e : (int)hipStreamSynchronize(S(stream));
}

int dil_sign_ctx_attempt_counts(dil_sign_ctx *c, uint32_t *h_counts, size_t n, void *stream) {
    if (!c || (n && !h_counts) || n > c->max_batch) return DIL_EINVAL;
    if (n == 0) return 0;
    const size_t off = with_params(c->mode, [](auto p) { return SignLayout<decltype(p)>::ATTEMPTS; });
    int e = copy2d(h_counts, sizeof(uint32_t), c->rows + off, c->row_pitch, sizeof(uint32_t), n,
                   hipMemcpyDeviceToHost, S(stream));
    return e ? e : (int)hipStreamSynchronize(S(stream));
}

int dil_sign_device(dil_sign_ctx *c, uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_sk, size_t sk_pitch,
                    const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, int engine,
                    void *stream) {
    return sign_enqueue(c, d_sig, sig_pitch, d_sk, sk_pitch, d_msg, d_moff, d_mlen, n, engine, S(stream), true);
}

}  // extern "C"

/* ------------------------------------------------------------ HIP graphs */

// A captured pipeline: every launch of one keygen / sign / verify call with
// its arguments fixed, replayed by one hipGraphLaunch.  Capture runs on a
// private stream in thread-local mode, so other threads' HIP calls do not
// interfere; the fixed sign stage's fork/join over the auxiliary streams
// becomes parallel branches of the graph.  No memset nodes are recorded
// (control words are cleared by a kernel).
struct dil_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t n = 0;
};

extern "C" int dil_graph_destroy(dil_graph *g) {
    if (!g) return 0;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return 0;
}

namespace {

template <class F>
int capture_graph(dil_graph **out, size_t n, F &&enqueue) {
    if (!out) return DIL_EINVAL;
    *out = nullptr;
    hipStream_t cap = nullptr;
    int e = (int)hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e) return e;
    dil_graph *g = new dil_graph();
    g->n = n;
    e = (int)hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (!e) {
        const int le = enqueue(cap);
        hipGraph_t gr = nullptr;
        const int ee = (int)hipStreamEndCapture(cap, &gr);
        g->graph = gr;
        e = le ? le : ee;
    }
    if (!e) e = (int)hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cap);
    if (e) {
        dil_graph_destroy(g);
        return e;
    }
    *out = g;
    return 0;
}

}  // namespace

extern "C" {

int dil_sign_graph_create(dil_graph **g, dil_sign_ctx *c, uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_sk,
                          size_t sk_pitch, const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                          size_t n) {
    if (!c || n == 0) return DIL_EINVAL;
    return capture_graph(g, n, [&](hipStream_t st) {
        return sign_enqueue(c, d_sig, sig_pitch, d_sk, sk_pitch, d_msg, d_moff, d_mlen, n, DIL_ENGINE_PERSISTENT,
                            st, false);
    });
}

int dil_verify_graph_create(dil_graph **g, int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch,
                            const uint8_t *d_pk, size_t pk_pitch, const uint8_t *d_msg, const uint64_t *d_moff,
                            const uint32_t *d_mlen, size_t n, uint8_t *d_work, size_t work_pitch) {
    if (n == 0) return DIL_EINVAL;
    return capture_graph(g, n, [&](hipStream_t st) {
        return dil_verify_device(mode, d_ret, d_sig, sig_pitch, d_pk, pk_pitch, d_msg, d_moff, d_mlen, n, d_work,
                                 work_pitch, st);
    });
}

int dil_keypair_graph_create(dil_graph **g, int mode, uint8_t *d_pk, size_t pk_pitch, uint8_t *d_sk,
                             size_t sk_pitch, const uint8_t *d_seeds, size_t seed_pitch, uint64_t rand_index,
                             size_t n, uint8_t *d_work, size_t work_pitch) {
    if (n == 0) return DIL_EINVAL;
    return capture_graph(g, n, [&](hipStream_t st) {
        return dil_keypair_device(mode, d_pk, pk_pitch, d_sk, sk_pitch, d_seeds, seed_pitch, rand_index, n, d_work,
                                  work_pitch, st);
    });
}

int dil_graph_launch(dil_graph *g, void *stream) {
    if (!g || !g->exec) return DIL_EINVAL;
    return (int)hipGraphLaunch(g->exec, S(stream));
}

size_t dil_graph_batch(const dil_graph *g) { return g ? g->n : 0; }

}  // extern "C"

/* ----------------------------------------------------------- key stores */

extern "C" {

size_t dil_keystore_row_bytes(int mode) {
    if (!valid_mode(mode)) return 0;
    return with_params(mode, [](auto p) { return KeyLayout<decltype(p)>::ROW; });
}

int dil_keystore_create(dil_keystore **out, int mode, size_t capacity) {
    if (!out || !valid_mode(mode) || capacity == 0 || capacity > 0xFFFFFFFFull) return DIL_EINVAL;
    dil_keystore *k = new dil_keystore();
    k->mode = mode;
    k->capacity = capacity;
    k->pitch = align256(dil_keystore_row_bytes(mode));
    k->has_sk.assign(capacity, 0);
    const int e = (int)hipMalloc((void **)&k->rows, k->pitch * capacity);
    if (e) {
        delete k;
        return e;
    }
    *out = k;
    return 0;
}

int dil_keystore_destroy(dil_keystore *k) {
    if (!k) return 0;
    (void)hipFree(k->rows);
    delete k;
    return 0;
}

size_t dil_keystore_capacity(const dil_keystore *k) { return k ? k->capacity : 0; }

int dil_keystore_mode(const dil_keystore *k) { return k ? k->mode : 0; }

int dil_keystore_has_secret(const dil_keystore *k, size_t idx) {
    return k && idx < k->capacity && k->has_sk[idx] ? 1 : 0;
}

int dil_keystore_load(dil_keystore *k, size_t first, const uint8_t *pk, size_t pk_pitch, const uint8_t *sk,
                      size_t sk_pitch, size_t n, void *stream) {
    if (!k || !pk) return DIL_EINVAL;
    if (n == 0) return 0;
    if (first > k->capacity || n > k->capacity - first) return DIL_ERANGE;
    if (pk_pitch < dil_public_key_bytes(k->mode) || (sk && sk_pitch < dil_secret_key_bytes(k->mode)))
        return DIL_EINVAL;
    const hipStream_t st = S(stream);
    return with_params(k->mode, [&](auto p) -> int {
        using P = decltype(p);
        using Z = Sizes<P>;
        using KY = KeyLayout<P>;
        uint8_t *row = k->rows + first * k->pitch;
        int e;
        if ((e = copy2d(row + KY::PK, k->pitch, pk, pk_pitch, Z::PK_BYTES, n,
                          caller_kind(pk, (n - 1) * pk_pitch + Z::PK_BYTES, hipMemcpyHostToDevice), st))) return e;
        if (sk && (e = copy2d(row + KY::SK, k->pitch, sk, sk_pitch, Z::SK_BYTES, n,
                                caller_kind(sk, (n - 1) * sk_pitch + Z::SK_BYTES, hipMemcpyHostToDevice), st)))
            return e;
        std::fill(k->has_sk.begin() + first, k->has_sk.begin() + first + n, sk ? 1 : 0);
        return launch_key_expand(k->mode, row, k->pitch, n, sk ? 1 : 0, st);
    });
}

int dil_keystore_keygen(dil_keystore *k, size_t first, const uint8_t *d_seeds, size_t seed_pitch,
                        uint64_t rand_index, size_t n, void *stream) {
    if (!k) return DIL_EINVAL;
    if (n == 0) return 0;
    if (first > k->capacity || n > k->capacity - first) return DIL_ERANGE;
    if (d_seeds && seed_pitch < 32) return DIL_EINVAL;
    const hipStream_t st = S(stream);
    return with_params(k->mode, [&](auto p) -> int {
        using KY = KeyLayout<decltype(p)>;
        uint8_t *row = k->rows + first * k->pitch;
        // the int area is keygen's scratch until the expansion overwrites it
        KeygenArgs a{row + KY::PK, k->pitch, row + KY::SK, k->pitch, row + KY::INTS, k->pitch,
                     d_seeds, seed_pitch, rand_index, n};
        int e = launch_keygen(k->mode, a, st);
        std::fill(k->has_sk.begin() + first, k->has_sk.begin() + first + n, 1);
        return e ? e : launch_key_expand(k->mode, row, k->pitch, n, 1, st);
    });
}

int dil_keystore_export(const dil_keystore *k, size_t first, size_t n, uint8_t *pk, size_t pk_pitch, uint8_t *sk,
                        size_t sk_pitch, void *stream) {
    if (!k) return DIL_EINVAL;
    if (n == 0) return 0;
    if (first > k->capacity || n > k->capacity - first) return DIL_ERANGE;
    if ((pk && pk_pitch < dil_public_key_bytes(k->mode)) || (sk && sk_pitch < dil_secret_key_bytes(k->mode)))
        return DIL_EINVAL;
    // a verify-only key has no secret key to export
    if (sk && std::find(k->has_sk.begin() + first, k->has_sk.begin() + first + n, 0) != k->has_sk.begin() + first + n)
        return DIL_EINVAL;
    const hipStream_t st = S(stream);
    return with_params(k->mode, [&](auto p) -> int {
        using P = decltype(p);
        using Z = Sizes<P>;
        using KY = KeyLayout<P>;
        const uint8_t *row = k->rows + first * k->pitch;
        int e;
        if (pk && (e = copy2d(pk, pk_pitch, row + KY::PK, k->pitch, Z::PK_BYTES, n,
                                caller_kind(pk, (n - 1) * pk_pitch + Z::PK_BYTES, hipMemcpyDeviceToHost), st))) return e;
        if (sk && (e = copy2d(sk, sk_pitch, row + KY::SK, k->pitch, Z::SK_BYTES, n,
                                caller_kind(sk, (n - 1) * sk_pitch + Z::SK_BYTES, hipMemcpyDeviceToHost), st))) return e;
        return 0;
    });
}

int dil_sign_device_resident(dil_sign_ctx *c, const dil_keystore *k, const uint32_t *d_key_idx, uint8_t *d_sig,
                             size_t sig_pitch, const uint8_t *d_msg, const uint64_t *d_moff,
                             const uint32_t *d_mlen, size_t n, int engine, void *stream) {
    if (!k) return DIL_EINVAL;
    return sign_enqueue(c, d_sig, sig_pitch, nullptr, 0, d_msg, d_moff, d_mlen, n, engine, S(stream), true, k,
                        d_key_idx);
}

int dil_verify_device_resident_ex(const dil_keystore *k, const uint32_t *d_key_idx, int32_t *d_ret,
                                  const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_msg,
                                  const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, uint8_t *d_work,
                                  size_t work_pitch, int flags, void *stream) {
    if (!k || !d_key_idx) return DIL_EINVAL;
    if (n == 0) return 0;
    if (flags & ~DIL_HASH_LONG_MESSAGES) return DIL_EINVAL;
    const int mode = k->mode;
    if (!d_ret || !d_sig || !d_work || !d_moff || !d_mlen || !d_msg) return DIL_EINVAL;
    if (sig_pitch < dil_signature_bytes(mode)) return DIL_EINVAL;
    if (work_pitch < dil_verify_work_row_bytes(mode) || (work_pitch & 255)) return DIL_ENOMEM;
    VerifyArgs a{d_sig, sig_pitch, k->rows, k->pitch, MsgRef{d_msg, d_moff, 0, d_mlen, 0}, d_work, work_pitch,
                 d_ret, 1, n, d_key_idx, k->rows, k->pitch, (uint32_t)k->capacity};
    a.row_mu = (flags & DIL_HASH_LONG_MESSAGES) ? 1 : 0;
    a.fork_mu = verify_fork_mu(n);
    return launch_verify(mode, a, S(stream));
}

int dil_verify_device_resident(const dil_keystore *k, const uint32_t *d_key_idx, int32_t *d_ret,
                               const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_msg,
                               const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, uint8_t *d_work,
                               size_t work_pitch, void *stream) {
    return dil_verify_device_resident_ex(k, d_key_idx, d_ret, d_sig, sig_pitch, d_msg, d_moff, d_mlen, n, d_work,
                                         work_pitch, 0, stream);
}

}  // extern "C"

/* ------------------------------------------------------ randomized sign */

extern "C" int dil_sign_device_randomized(dil_sign_ctx *c, const dil_keystore *ks, const uint32_t *d_key_idx,
                                          uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_sk, size_t sk_pitch,
                                          const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                                          const uint8_t *d_rhoprime, size_t rhoprime_pitch, size_t n, int engine,
                                          void *stream) {
    if (!d_rhoprime || rhoprime_pitch < 64) return DIL_EINVAL;
    if (ks && !d_key_idx) return DIL_EINVAL;
    return sign_enqueue(c, d_sig, sig_pitch, d_sk, sk_pitch, d_msg, d_moff, d_mlen, n, engine, S(stream), true, ks,
                        d_key_idx, d_rhoprime, rhoprime_pitch);
}
