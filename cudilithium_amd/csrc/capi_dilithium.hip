// C-ABI entry points of the keygen / sign / verify pipelines: the drop-in
// replacements of the reference's api.cuh surface and the device-resident
// extensions.
#include <hip/hip_runtime.h>

#include <cstring>

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
        (void)engine;
        e = sign_perstage(mode, a, item_slots, temp_slots, exec_threshold,
                          reinterpret_cast<SignLutEntry *>(lut->sign_lut), lut->h_exec_lut, lut->d_exec_lut,
                          lut->h_done_lut, lut->d_done_lut, reinterpret_cast<uint64_t *>(lut->h_copy_lut),
                          reinterpret_cast<uint64_t *>(lut->d_copy_lut), st);
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
