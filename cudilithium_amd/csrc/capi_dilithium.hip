// C-ABI entry points of the keygen / sign / verify pipelines: the drop-in
// replacements of the reference's api.cuh surface and the device-resident
// extensions.
#include <hip/hip_runtime.h>

#include <cstring>
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

int copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height, hipMemcpyKind k,
           hipStream_t st) {
    if (width == 0 || height == 0) return 0;
    return (int)hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, k, st);
}

}  // namespace

static_assert(sizeof(dil_sign_lut_element) == sizeof(SignLutEntry), "sign_lut_element layout");
static_assert(sizeof(dil_copy_lut_element) == 2 * sizeof(uint64_t), "copy_lut_element layout");

extern "C" {

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

static int sign_dropin(int mode, int engine, uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                       size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool, size_t sign_pitch,
                       uint8_t *d_temp_pool, size_t temp_pitch, dil_task_lut *lut, size_t exec_threshold,
                       size_t batch, void *stream) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (siglen) *siglen = dil_signature_bytes(mode);
    if (batch == 0) return 0;
    if (!sig || !sk || !d_sign_pool || !lut || exec_threshold == 0 || (mlen && !m)) return DIL_EINVAL;
    if (batch > 65536) return DIL_ERANGE;  // (item << 16 | nonce) entries, api.cu:224
    if (sign_pitch < dil_sign_pool_row_bytes(mode, mlen)) return DIL_ENOMEM;
    if (!d_temp_pool || temp_pitch < dil_sign_temp_row_bytes(mode)) return DIL_ENOMEM;
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
        const SlotRef item_slots{d_sign_pool + SLy::SLOT, sign_pitch, d_sign_pool + SLy::SIG, sign_pitch};
        const SlotRef temp_slots{d_temp_pool + TempLayout<P>::SLOT, temp_pitch, d_temp_pool + TempLayout<P>::SIG,
                                 temp_pitch};
        if (engine == DIL_ENGINE_PERSTAGE) {
            e = sign_perstage(mode, a, item_slots, temp_slots, exec_threshold,
                              reinterpret_cast<SignLutEntry *>(lut->sign_lut), lut->h_exec_lut, lut->d_exec_lut,
                              lut->h_done_lut, lut->d_done_lut, reinterpret_cast<uint64_t *>(lut->h_copy_lut),
                              reinterpret_cast<uint64_t *>(lut->d_copy_lut), st);
        } else {
            // the exec_threshold temp rows are the attempt slots; the first
            // word of d_copy_lut is the work-queue head
            if ((e = launch_sign_fixed(mode, a, st))) return e;
            uint32_t *ctl = reinterpret_cast<uint32_t *>(lut->d_copy_lut);
            e = launch_sign_persistent(mode, a, temp_slots, exec_threshold, default_slots_per_block(exec_threshold),
                                       0, 0, ctl, st);
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
        if (pitch % 4) return DIL_EINVAL;
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

int dil_verify_device(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_pk,
                      size_t pk_pitch, const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                      size_t n, uint8_t *d_work, size_t work_pitch, void *stream) {
    if (!valid_mode(mode)) return DIL_EINVAL;
    if (n == 0) return 0;
    if (!d_ret || !d_sig || !d_pk || !d_work || !d_moff || !d_mlen || !d_msg) return DIL_EINVAL;
    if (sig_pitch < dil_signature_bytes(mode) || pk_pitch < dil_public_key_bytes(mode)) return DIL_EINVAL;
    if (work_pitch < dil_verify_work_row_bytes(mode) || (work_pitch & 255)) return DIL_ENOMEM;
    VerifyArgs a{d_sig, sig_pitch, d_pk, pk_pitch, MsgRef{d_msg, d_moff, 0, d_mlen, 0}, d_work, work_pitch, d_ret, 1, n};
    return launch_verify(mode, a, S(stream));
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
};

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
    if (!out || !valid_mode(mode) || max_batch == 0 || nslots == 0) return DIL_EINVAL;
    dil_sign_ctx *c = new dil_sign_ctx();
    c->mode = mode;
    c->max_batch = max_batch;
    c->nslots = nslots;
    c->row_pitch = align256(dil_sign_pool_row_bytes(mode, 0));
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
    if (!c) return DIL_EINVAL;
    if (n == 0) return 0;
    if (n > c->max_batch) return DIL_ERANGE;
    if (!d_sig || !d_sk || !d_msg || !d_moff || !d_mlen) return DIL_EINVAL;
    if (sig_pitch < dil_signature_bytes(c->mode) || sk_pitch < dil_secret_key_bytes(c->mode)) return DIL_EINVAL;
    if (engine != DIL_ENGINE_PERSISTENT && engine != DIL_ENGINE_PERSTAGE) return DIL_EINVAL;
    if (engine == DIL_ENGINE_PERSTAGE && n > 65536) return DIL_ERANGE;
    const hipStream_t st = S(stream);
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
        const SlotRef slots{c->slots + TempLayout<P>::SLOT, c->slot_pitch, c->slots + TempLayout<P>::SIG,
                            c->slot_pitch};
        int e;
        hipEvent_t *e3 = c->timed < c->timing_cap ? &c->ev[3 * c->timed] : nullptr;
        if (e3 && (e = (int)hipEventRecord(e3[0], st))) return e;
        if (engine == DIL_ENGINE_PERSTAGE) {
            // the per-stage scheduler launches the fixed part itself
            if (e3 && (e = (int)hipEventRecord(e3[1], st))) return e;
            const SlotRef items{c->rows + SignLayout<P>::SLOT, c->row_pitch, d_sig, sig_pitch};
            e = sign_perstage(c->mode, a, items, slots, c->nslots, c->sign_lut.data(), c->h_exec, c->d_exec,
                              c->h_done, c->d_done, c->h_copy, c->d_copy, st);
        } else {
            if ((e = launch_sign_fixed(c->mode, a, st))) return e;
            if (e3 && (e = (int)hipEventRecord(e3[1], st))) return e;
            const int spb = c->slots_per_block ? c->slots_per_block : default_slots_per_block(c->nslots);
            e = launch_sign_persistent(c->mode, a, slots, c->nslots, spb, c->occupancy, c->blocks, c->d_head, st);
            if (e == -3) return DIL_ERANGE;
        }
        if (e) return e;
        if (e3) {
            if ((e = (int)hipEventRecord(e3[2], st))) return e;
            ++c->timed;
        }
        return 0;
    });
}

}  // extern "C"
