/*
 * cudilithium_amd: batched CRYSTALS-Dilithium (round-3 v3.1; modes 2, 3, 5)
 * keygen / deterministic sign / verify on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libcudilithium_amd.so.  Plain pointers and sizes only; streams are
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Three groups of entry points:
 *   1. Drop-in replacements of the reference's batched API
 *      (include/api.cuh:27-63 of encryptorion-lab/cuDilithium): same
 *      argument order, meaning, pool ownership and return conventions.  The
 *      reference builds one library per mode with identical symbol names; here
 *      one library exports dil2_*, dil3_* and dil5_* (and dil_* taking the
 *      mode as the first argument).  include/cudilithium_amd/api.hpp maps the
 *      reference's C++ names onto these under -DDILITHIUM_MODE.
 *   2. Extensions: explicit-seed keygen, device-resident batches with
 *      per-item message lengths, and the standalone NTT.
 *   3. Runtime helpers (device memory, copies, streams, events) so that a
 *      host binding needs no other GPU runtime.
 *
 * Return values: 0 on success; a positive hipError_t value when a HIP call
 * fails; a negative DIL_E* code for invalid arguments.
 */
#ifndef CUDILITHIUM_AMD_H
#define CUDILITHIUM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIL_OK 0
#define DIL_EINVAL (-1)   /* bad mode, size or pointer */
#define DIL_ENOMEM (-2)   /* a caller-provided buffer is too small */
#define DIL_ERANGE (-3)   /* a size exceeds what the engine supports */

/* ---------------------------------------------------------------------- */
/* Parameters                                                             */
/* ---------------------------------------------------------------------- */

/* CRYPTO_PUBLICKEYBYTES / CRYPTO_SECRETKEYBYTES / CRYPTO_BYTES of a mode
 * (include/params.h:71-73); 0 for an unknown mode. */
size_t dil_public_key_bytes(int mode);
size_t dil_secret_key_bytes(int mode);
size_t dil_signature_bytes(int mode);

/* ---------------------------------------------------------------------- */
/* Standalone NTT (include/ntt.cuh:30-62, src/ntt.cu:77-207)              */
/* Polynomials are 256 int32 coefficients; consecutive polynomials start  */
/* pitch_words int32 apart (pitch_words >= 256).  Device pointers.        */
/* Results are int32 bit-exact with the reference's radix-8 ntt_inner /   */
/* invntt_inner.                                                          */
/* ---------------------------------------------------------------------- */

/* In-place forward NTT of npoly polynomials (replaces ntt_radix8_kernel). */
int dil_ntt_forward(int32_t *d_poly, size_t npoly, size_t pitch_words, void *stream);
/* In-place inverse NTT with Montgomery factor (replaces intt_radix8_kernel). */
int dil_ntt_inverse(int32_t *d_poly, size_t npoly, size_t pitch_words, void *stream);
/* d_c = INTT(NTT(d_a) o_mont d_b), d_b already in the NTT domain; fused in
 * one launch.  d_c may alias d_a. */
int dil_ntt_pointwise_pipeline(int32_t *d_c, const int32_t *d_a, const int32_t *d_b, size_t npoly,
                               size_t pitch_words, void *stream);

/* Keccak-f[1600] applied to n states stored lane-major (word j of state i at
 * d_states[j * n + i]).  Device pointer. */
int dil_keccak_f1600_batch(uint64_t *d_states, size_t n, void *stream);

/* ---------------------------------------------------------------------- */
/* Runtime helpers                                                        */
/* ---------------------------------------------------------------------- */
int dil_device_count(int *count);
int dil_set_device(int device);
int dil_device_name(char *buf, size_t len);
int dil_device_cu_count(void);
int dil_malloc_device(void **ptr, size_t bytes);
int dil_malloc_pitch(void **ptr, size_t *pitch, size_t width, size_t height);
int dil_free_device(void *ptr);
int dil_malloc_host(void **ptr, size_t bytes); /* pinned */
int dil_free_host(void *ptr);
int dil_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int dil_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int dil_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int dil_memset_device(void *dst, int value, size_t bytes, void *stream);
int dil_stream_create(void **stream);
int dil_stream_destroy(void *stream);
int dil_stream_synchronize(void *stream);
int dil_device_synchronize(void);
int dil_event_create(void **event);
int dil_event_destroy(void *event);
int dil_event_record(void *event, void *stream);
int dil_event_synchronize(void *event);
int dil_event_elapsed_ms(float *ms, void *start, void *stop);
const char *dil_error_string(int code);

#ifdef __cplusplus
}
#endif

#endif /* CUDILITHIUM_AMD_H */
