// Source-level drop-in for the reference's C++ interface
// (encryptorion-lab/cuDilithium include/api.cuh:17-63 and include/params.h):
// the same function names, parameter lists, default arguments, task_lut
// shape and size macros, selected at compile time with -DDILITHIUM_MODE=2|3|5
// as in the reference (params.h:14-47), forwarding to libcudilithium_amd.so.
//
// Differences a caller can see: cudaStream_t is hipStream_t; the functions
// return the library's status codes (0 on success) where the reference always
// returns 0; crypto_sign_verify returns that status too (void in the
// reference; ignoring it is source compatible).
//
// -DDILITHIUM_RANDOMIZED_SIGNING selects randomized signing as in the
// reference (api.cu:143-148): rhoprime comes from the library's random source
// (getrandom(2) unless the caller installs its own function with
// dil_set_randombytes, the counterpart of linking the reference's randombytes,
// benches/trng.c:21-80).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../cudilithium_amd.h"

#ifndef DILITHIUM_MODE
#define DILITHIUM_MODE 2
#endif

#define SEEDBYTES 32
#define CRHBYTES 64
#define DILITHIUM_N 256
#define DILITHIUM_Q 8380417
#define DILITHIUM_D 13
#define ROOT_OF_UNITY 1753

#if DILITHIUM_MODE == 2
#define DILITHIUM_K 4
#define DILITHIUM_L 4
#define ETA 2
#define TAU 39
#define BETA 78
#define GAMMA1 (1 << 17)
#define GAMMA2 ((DILITHIUM_Q - 1) / 88)
#define OMEGA 80
#elif DILITHIUM_MODE == 3
#define DILITHIUM_K 6
#define DILITHIUM_L 5
#define ETA 4
#define TAU 49
#define BETA 196
#define GAMMA1 (1 << 19)
#define GAMMA2 ((DILITHIUM_Q - 1) / 32)
#define OMEGA 55
#elif DILITHIUM_MODE == 5
#define DILITHIUM_K 8
#define DILITHIUM_L 7
#define ETA 2
#define TAU 60
#define BETA 120
#define GAMMA1 (1 << 19)
#define GAMMA2 ((DILITHIUM_Q - 1) / 32)
#define OMEGA 75
#else
#error "DILITHIUM_MODE must be 2, 3 or 5"
#endif

#define POLYT1_PACKEDBYTES 320
#define POLYT0_PACKEDBYTES 416
#define POLYVECH_PACKEDBYTES (OMEGA + DILITHIUM_K)
#if GAMMA1 == (1 << 17)
#define POLYZ_PACKEDBYTES 576
#else
#define POLYZ_PACKEDBYTES 640
#endif
#if GAMMA2 == (DILITHIUM_Q - 1) / 88
#define POLYW1_PACKEDBYTES 192
#else
#define POLYW1_PACKEDBYTES 128
#endif
#if ETA == 2
#define POLYETA_PACKEDBYTES 96
#else
#define POLYETA_PACKEDBYTES 128
#endif

#define CRYPTO_PUBLICKEYBYTES (SEEDBYTES + DILITHIUM_K * POLYT1_PACKEDBYTES)
#define CRYPTO_SECRETKEYBYTES \
    (3 * SEEDBYTES + DILITHIUM_L * POLYETA_PACKEDBYTES + DILITHIUM_K * POLYETA_PACKEDBYTES + DILITHIUM_K * POLYT0_PACKEDBYTES)
#define CRYPTO_BYTES (SEEDBYTES + DILITHIUM_L * POLYZ_PACKEDBYTES + POLYVECH_PACKEDBYTES)

#define ALIGN_TO_256_BYTES(x) ((((x) + 255) / 256) * 256)

struct sign_lut_element {
    int done_flag = 0;
    uint16_t nonce = 0;
};

struct copy_lut_element {
    size_t exec_idx;
    size_t sign_idx;
};

struct task_lut {
    std::vector<sign_lut_element> sign_lut;
    uint32_t *h_exec_lut;
    uint8_t *h_done_lut;
    copy_lut_element *h_copy_lut;
    uint32_t *d_exec_lut;
    uint8_t *d_done_lut;
    copy_lut_element *d_copy_lut;
};

static_assert(sizeof(sign_lut_element) == sizeof(dil_sign_lut_element), "sign_lut_element layout");
static_assert(sizeof(copy_lut_element) == sizeof(dil_copy_lut_element), "copy_lut_element layout");

inline int crypto_sign_keypair(uint8_t *pk, uint8_t *sk, uint8_t *d_keypair_mem_pool, size_t keypair_mem_pool_pitch,
                               size_t batch_size = 1, hipStream_t stream = nullptr, size_t rand_index = 0) {
    return dil_crypto_sign_keypair(DILITHIUM_MODE, pk, sk, d_keypair_mem_pool, keypair_mem_pool_pitch, batch_size,
                                   stream, rand_index);
}

inline int crypto_sign_signature(uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m, size_t m_pitch,
                                 size_t mlen, const uint8_t *sk, uint8_t *d_sign_mem_pool,
                                 size_t d_sign_mem_pool_pitch, uint8_t *d_temp_mem_pool,
                                 size_t d_temp_mem_pool_pitch, task_lut &lut, size_t exec_threshold,
                                 size_t batch_size = 1, hipStream_t stream = nullptr) {
    if (lut.sign_lut.size() < batch_size) lut.sign_lut.resize(batch_size);
    dil_task_lut c{reinterpret_cast<dil_sign_lut_element *>(lut.sign_lut.data()), lut.sign_lut.size(),
                   lut.h_exec_lut, lut.h_done_lut, reinterpret_cast<dil_copy_lut_element *>(lut.h_copy_lut),
                   lut.d_exec_lut, lut.d_done_lut, reinterpret_cast<dil_copy_lut_element *>(lut.d_copy_lut)};
#ifdef DILITHIUM_RANDOMIZED_SIGNING
    return dil_crypto_sign_signature_randomized(DILITHIUM_MODE, sig, sig_pitch, siglen, m, m_pitch, mlen, sk,
                                                d_sign_mem_pool, d_sign_mem_pool_pitch, d_temp_mem_pool,
                                                d_temp_mem_pool_pitch, &c, exec_threshold, batch_size, stream);
#else
    return dil_crypto_sign_signature(DILITHIUM_MODE, sig, sig_pitch, siglen, m, m_pitch, mlen, sk, d_sign_mem_pool,
                                     d_sign_mem_pool_pitch, d_temp_mem_pool, d_temp_mem_pool_pitch, &c,
                                     exec_threshold, batch_size, stream);
#endif
}

inline int crypto_sign_verify(int *ret, const uint8_t *sig, size_t sig_pitch, size_t siglen, const uint8_t *m,
                              size_t m_pitch, size_t mlen, const uint8_t *pk, uint8_t *d_verify_mem_pool,
                              size_t verify_mem_pool_pitch, size_t batch_size = 1, hipStream_t stream = nullptr) {
    return dil_crypto_sign_verify(DILITHIUM_MODE, ret, sig, sig_pitch, siglen, m, m_pitch, mlen, pk,
                                  d_verify_mem_pool, verify_mem_pool_pitch, batch_size, stream);
}
