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
/* Drop-in batched API (include/api.cuh:27-63)                            */
/* ---------------------------------------------------------------------- */

/* sign_lut_element (api.cuh:31-34) */
typedef struct {
    int done_flag;
    uint16_t nonce;
} dil_sign_lut_element;

/* copy_lut_element (api.cuh:36-39) */
typedef struct {
    uint64_t exec_idx;
    uint64_t sign_idx;
} dil_copy_lut_element;

/* task_lut (api.cuh:41-49).  The reference keeps sign_lut in a std::vector;
 * here the caller passes an array of at least `batch` elements.  All other
 * members are allocated by the caller exactly as in the reference
 * (tests/test_vectors.cu:104-113): exec_threshold entries each, h_* pinned
 * host memory, d_* device memory. */
typedef struct {
    dil_sign_lut_element *sign_lut;
    size_t sign_lut_len;
    uint32_t *h_exec_lut;
    uint8_t *h_done_lut;
    dil_copy_lut_element *h_copy_lut;
    uint32_t *d_exec_lut;
    uint8_t *d_done_lut;
    dil_copy_lut_element *d_copy_lut;
} dil_task_lut;

#define DIL_POOL_KEYPAIR 0
#define DIL_POOL_SIGN 1
#define DIL_POOL_SIGN_TEMP 2
#define DIL_POOL_VERIFY 3

#define DIL_ENGINE_PERSISTENT 0 /* device work queue, one persistent launch */
#define DIL_ENGINE_PERSTAGE 1   /* one launch per stage, host-driven rounds */

/* Minimum pitch of each pool row this library needs.  Each is at most the
 * row size the reference's callers compute (tests/test_vectors.cu:64-70,
 * 81-100, 117-128, 149-166), which dil_reference_pool_row_bytes returns, so
 * pools sized by the reference's formulas work unchanged. */
size_t dil_keypair_pool_row_bytes(int mode);
size_t dil_sign_pool_row_bytes(int mode, size_t mlen);
size_t dil_sign_temp_row_bytes(int mode);
size_t dil_verify_pool_row_bytes(int mode, size_t mlen);
size_t dil_reference_pool_row_bytes(int mode, int which, size_t mlen);

/* Host buffers of the three drop-in calls may be pinned (hipHostMalloc) or
 * pageable.  Pinned rows are copied by a kernel on `stream` through their
 * device mapping, asynchronously; pageable rows go through a 32 MiB pinned
 * staging buffer that the library allocates on first use and keeps until
 * dil_release_host_staging or the end of the process (packed by up to 4
 * host threads), and those copies complete before the call returns. */
/* Frees the pinned staging buffer of pageable copies (after any copy that
 * holds it); the next pageable copy allocates it again.  0 or hipError_t. */
int dil_release_host_staging(void);

/* crypto_sign_keypair (api.cuh:27-29, api.cu:26-61): batch keys whose seeds
 * are SHAKE128(le64(rand_index + 2i + 1))[0:32]; pk/sk are host buffers
 * (packed, stride CRYPTO_PUBLICKEYBYTES / CRYPTO_SECRETKEYBYTES), copied back
 * asynchronously on `stream` (the caller synchronises). */
int dil_crypto_sign_keypair(int mode, uint8_t *pk, uint8_t *sk, uint8_t *d_keypair_mem_pool,
                            size_t keypair_mem_pool_pitch, size_t batch_size, void *stream,
                            size_t rand_index);

/* crypto_sign_signature (api.cuh:51-56, api.cu:63-452): deterministic
 * signatures of batch messages of mlen bytes (host, stride m_pitch) under
 * host secret keys; signatures to host `sig` with stride sig_pitch;
 * *siglen = CRYPTO_BYTES.  Per item the result is the attempt with the first
 * successful nonce, as in the reference.  This entry point runs the
 * persistent work-queue engine.  Attempt slots: the exec_threshold temp
 * rows when batch_size <= exec_threshold; otherwise up to batch_size slots,
 * each with its state in a sign-pool row and its candidate signature in the
 * temp pool read as one array.  batch_size <= 65536. */
int dil_crypto_sign_signature(int mode, uint8_t *sig, size_t sig_pitch, size_t *siglen,
                              const uint8_t *m, size_t m_pitch, size_t mlen, const uint8_t *sk,
                              uint8_t *d_sign_mem_pool, size_t d_sign_mem_pool_pitch,
                              uint8_t *d_temp_mem_pool, size_t d_temp_mem_pool_pitch,
                              dil_task_lut *lut, size_t exec_threshold, size_t batch_size,
                              void *stream);
/* Same contract, run as the reference architecture: one launch per stage and
 * host-driven rejection rounds of exec_threshold attempts. */
int dil_crypto_sign_signature_perstage(int mode, uint8_t *sig, size_t sig_pitch, size_t *siglen,
                                       const uint8_t *m, size_t m_pitch, size_t mlen,
                                       const uint8_t *sk, uint8_t *d_sign_mem_pool,
                                       size_t d_sign_mem_pool_pitch, uint8_t *d_temp_mem_pool,
                                       size_t d_temp_mem_pool_pitch, dil_task_lut *lut,
                                       size_t exec_threshold, size_t batch_size, void *stream);

/* crypto_sign_signature as built with -DDILITHIUM_RANDOMIZED_SIGNING
 * (api.cu:143-148): per item, rhoprime is 64 bytes drawn from the random
 * source (one draw of 64 * batch_size bytes, item i taking bytes
 * [64 i, 64 i + 64)) instead of SHAKE256(key || mu); otherwise the same
 * contract as dil_crypto_sign_signature (persistent engine).  Returns the
 * random source's error, if it reports one, before any GPU work. */
int dil_crypto_sign_signature_randomized(int mode, uint8_t *sig, size_t sig_pitch, size_t *siglen,
                                         const uint8_t *m, size_t m_pitch, size_t mlen,
                                         const uint8_t *sk, uint8_t *d_sign_mem_pool,
                                         size_t d_sign_mem_pool_pitch, uint8_t *d_temp_mem_pool,
                                         size_t d_temp_mem_pool_pitch, dil_task_lut *lut,
                                         size_t exec_threshold, size_t batch_size, void *stream);
/* The random source of randomized signing (the reference's randombytes,
 * benches/trng.c:21-80, which reads the OS generator).  The default reads
 * getrandom(2); dil_set_randombytes installs the caller's function (NULL
 * restores the default).  A source returns 0 on success.  Process-wide. */
typedef int (*dil_randombytes_fn)(uint8_t *out, size_t outlen);
void dil_set_randombytes(dil_randombytes_fn fn);
int dil_randombytes(uint8_t *out, size_t outlen);

/* crypto_sign_verify (api.cuh:58-63, api.cu:454-515): ret[i] = 0 for a valid
 * signature, -1 otherwise; siglen != CRYPTO_BYTES sets every ret to -1
 * without GPU work.  ret is copied back asynchronously on `stream`. */
int dil_crypto_sign_verify(int mode, int *ret, const uint8_t *sig, size_t sig_pitch, size_t siglen,
                           const uint8_t *m, size_t m_pitch, size_t mlen, const uint8_t *pk,
                           uint8_t *d_verify_mem_pool, size_t verify_mem_pool_pitch,
                           size_t batch_size, void *stream);

/* Per-mode names with the reference's exact parameter lists (one library per
 * mode in the reference, src/CMakeLists.txt:3-16). */
int dil2_crypto_sign_keypair(uint8_t *pk, uint8_t *sk, uint8_t *d_pool, size_t pitch, size_t batch,
                             void *stream, size_t rand_index);
int dil3_crypto_sign_keypair(uint8_t *pk, uint8_t *sk, uint8_t *d_pool, size_t pitch, size_t batch,
                             void *stream, size_t rand_index);
int dil5_crypto_sign_keypair(uint8_t *pk, uint8_t *sk, uint8_t *d_pool, size_t pitch, size_t batch,
                             void *stream, size_t rand_index);
int dil2_crypto_sign_signature(uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                               size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool,
                               size_t sign_pitch, uint8_t *d_temp_pool, size_t temp_pitch,
                               dil_task_lut *lut, size_t exec_threshold, size_t batch, void *stream);
int dil3_crypto_sign_signature(uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                               size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool,
                               size_t sign_pitch, uint8_t *d_temp_pool, size_t temp_pitch,
                               dil_task_lut *lut, size_t exec_threshold, size_t batch, void *stream);
int dil5_crypto_sign_signature(uint8_t *sig, size_t sig_pitch, size_t *siglen, const uint8_t *m,
                               size_t m_pitch, size_t mlen, const uint8_t *sk, uint8_t *d_sign_pool,
                               size_t sign_pitch, uint8_t *d_temp_pool, size_t temp_pitch,
                               dil_task_lut *lut, size_t exec_threshold, size_t batch, void *stream);
int dil2_crypto_sign_verify(int *ret, const uint8_t *sig, size_t sig_pitch, size_t siglen,
                            const uint8_t *m, size_t m_pitch, size_t mlen, const uint8_t *pk,
                            uint8_t *d_pool, size_t pitch, size_t batch, void *stream);
int dil3_crypto_sign_verify(int *ret, const uint8_t *sig, size_t sig_pitch, size_t siglen,
                            const uint8_t *m, size_t m_pitch, size_t mlen, const uint8_t *pk,
                            uint8_t *d_pool, size_t pitch, size_t batch, void *stream);
int dil5_crypto_sign_verify(int *ret, const uint8_t *sig, size_t sig_pitch, size_t siglen,
                            const uint8_t *m, size_t m_pitch, size_t mlen, const uint8_t *pk,
                            uint8_t *d_pool, size_t pitch, size_t batch, void *stream);

/* ---------------------------------------------------------------------- */
/* Device-resident batches (extensions; all pointers are device memory)   */
/* Item i of an array lives at base + i * pitch.  Messages: item i is      */
/* d_msg + d_moff[i], d_mlen[i] bytes (per-item lengths).                 */
/* ---------------------------------------------------------------------- */

/* Keygen from explicit 32-byte seeds (d_seeds != NULL), or from the
 * reference's counter seeds (d_seeds == NULL, rand_index as above).
 * d_work: n rows of at least dil_keypair_work_row_bytes(mode), pitch a
 * multiple of 8. */
size_t dil_keypair_work_row_bytes(int mode);
int dil_keypair_device(int mode, uint8_t *d_pk, size_t pk_pitch, uint8_t *d_sk, size_t sk_pitch,
                       const uint8_t *d_seeds, size_t seed_pitch, uint64_t rand_index, size_t n,
                       uint8_t *d_work, size_t work_pitch, void *stream);

/* Verify n (sig, msg, pk) triples; d_ret[i] = 0 or -1.  d_work: n rows of at
 * least dil_verify_work_row_bytes(mode), pitch a multiple of 256. */
size_t dil_verify_work_row_bytes(int mode);
int dil_verify_device(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch,
                      const uint8_t *d_pk, size_t pk_pitch, const uint8_t *d_msg,
                      const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, uint8_t *d_work,
                      size_t work_pitch, void *stream);
/* Message hashing of the verify / sign calls below: by default one lane
 * absorbs a message serially (one SHAKE256 permutation per 136 bytes, the
 * throughput form).  DIL_HASH_LONG_MESSAGES gives each message a 16-lane
 * row of ten active lanes instead, which shortens the serial chain about
 * 2.5 times at more instructions per message: for batches whose latency
 * is set by long messages.  Results are identical. */
#define DIL_HASH_LONG_MESSAGES 1
int dil_verify_device_ex(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch,
                         const uint8_t *d_pk, size_t pk_pitch, const uint8_t *d_msg,
                         const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, uint8_t *d_work,
                         size_t work_pitch, int flags, void *stream);

/* dil_verify_device that allocates nothing: no counterpart in the reference
 * (whose callers pass pitched pools); for callers that cannot allocate device
 * memory -- code under HIP-graph capture, serving loops that forbid hipMalloc.
 * The work rows live in a fixed array of DIL_VERIFY_NOALLOC_SCRATCH_BYTES per
 * GPU inside the library (held until the process ends), so n is limited to
 * dil_verify_noalloc_max_items(mode) (8192 for mode 5, 9362 for modes 2 and
 * 3); larger n gives DIL_ERANGE.  All calls on one GPU share that array:
 * calls that overlap on one GPU race, so serialise them (one stream, or
 * synchronise between calls).  Results equal dil_verify_device's. */
#define DIL_VERIFY_NOALLOC_SCRATCH_BYTES (16u << 20)
size_t dil_verify_noalloc_max_items(int mode);
int dil_verify_device_noalloc(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch,
                              const uint8_t *d_pk, size_t pk_pitch, const uint8_t *d_msg,
                              const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, void *stream);
/* dil_verify_device with its three launches (vf_prep, vf_main, vf_final)
 * timed by HIP events on `stream`; waits for the call and writes the three
 * durations in ms to kernel_ms3.  A measurement aid, same results. */
int dil_verify_device_timed(int mode, int32_t *d_ret, const uint8_t *d_sig, size_t sig_pitch,
                            const uint8_t *d_pk, size_t pk_pitch, const uint8_t *d_msg,
                            const uint64_t *d_moff, const uint32_t *d_mlen, size_t n,
                            uint8_t *d_work, size_t work_pitch, float *kernel_ms3, void *stream);

/* Run-time tilings of the batched kernels, as the autotuner
 * (cudilithium_amd/autotune.py) records them: per kind, mode and batch
 * bucket (0: up to 1,024 items, 1: 4,096, 2: 16,384, 3: more; -1 = all
 * buckets) a value, 0 restoring the compiled default.  Kinds: keys per
 * workgroup of keygen's main kernel and items per workgroup of verify's
 * (2, 4, 8 or 16 where its LDS fits: defaults 16 / 8 / 4 for modes 2 / 3 / 5),
 * wavefronts per workgroup of the standalone NTT (2, 4 or 8; default 4; mode
 * ignored).  Process-wide; results never depend on them.  DIL_EINVAL for a
 * value the kernels are not built for. */
#define DIL_TILE_KEYGEN_ITEMS 0
#define DIL_TILE_VERIFY_ITEMS 1
#define DIL_TILE_NTT_WAVES 2
int dil_set_kernel_tiling(int kind, int mode, int bucket, int value);
/* The value in effect for a batch of n items (0: the compiled default). */
int dil_get_kernel_tiling(int kind, int mode, size_t n);

/* Signing context: owns the per-item rows (dil_sign_pool_row_bytes(mode, 0)
 * each), nslots attempt slots and the scheduler tables of either engine. */
typedef struct dil_sign_ctx dil_sign_ctx;
int dil_sign_ctx_create(dil_sign_ctx **ctx, int mode, size_t max_batch, size_t nslots);
int dil_sign_ctx_destroy(dil_sign_ctx *ctx);
/* Persistent engine tiling: slots per workgroup (0 = automatic, or 1, 8, 16,
 * 32), register budget as waves per SIMD (0 = automatic, or 2, 3, 4) and
 * workgroup count (0 = nslots / slots_per_block). */
int dil_sign_ctx_set_tiling(dil_sign_ctx *ctx, int slots_per_block, int occupancy, int blocks);
/* Persistent engine: at most items_per_block items owned by a workgroup at a
 * time (0 = slots per workgroup).  Below that, each owned item gets several
 * slots (consecutive nonces) per round, so small batches finish in fewer
 * rounds at the price of speculative attempts; results are unchanged.
 * OR DIL_CLAIM_GUIDED into items_per_block for guided claims: once the queue
 * holds fewer than items_per_block items per resident workgroup, a workgroup
 * owns at most (remaining items / resident workgroups), so the batch's last
 * items spread over all workgroups (more speculation, shorter tail). */
#define DIL_CLAIM_GUIDED 0x10000
int dil_sign_ctx_set_claim(dil_sign_ctx *ctx, int items_per_block);
/* Records HIP events on the call's stream around the fixed stage and the
 * rejection loop of the next max_calls dil_sign_device calls (0 turns timing
 * off and forgets earlier records).  dil_sign_ctx_timing waits for call
 * `call` (0-based) and returns its two durations. */
int dil_sign_ctx_enable_timing(dil_sign_ctx *ctx, int max_calls);
int dil_sign_ctx_timed_calls(dil_sign_ctx *ctx);
int dil_sign_ctx_timing(dil_sign_ctx *ctx, int call, float *fixed_ms, float *loop_ms);
/* Deterministic signatures of n messages (per-item lengths) under n secret
 * keys, all in device memory; engine DIL_ENGINE_PERSISTENT or
 * DIL_ENGINE_PERSTAGE (n <= 65536).  Asynchronous for the persistent engine;
 * the per-stage engine synchronises `stream` once per rejection round. */
int dil_sign_device(dil_sign_ctx *ctx, uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_sk,
                    size_t sk_pitch, const uint8_t *d_msg, const uint64_t *d_moff,
                    const uint32_t *d_mlen, size_t n, int engine, void *stream);
/* After a persistent-engine call: per item, the number of attempts the
 * sequential specification makes (winning nonce + 1), copied to host and
 * synchronised on `stream`. */
int dil_sign_ctx_attempt_counts(dil_sign_ctx *ctx, uint32_t *h_counts, size_t n, void *stream);
/* Status of the last call (synchronises `stream`): 0, or a combination of
 * DIL_SIGN_STATUS_ROUND_CAP (a persistent workgroup stopped at its round
 * limit; never expected; the batch's signatures are then incomplete) and
 * DIL_SIGN_STATUS_NO_SECRET (resident signing: an item's key is verify-only). */
#define DIL_SIGN_STATUS_ROUND_CAP 1
#define DIL_SIGN_STATUS_NO_SECRET 2
int dil_sign_ctx_status(dil_sign_ctx *ctx, int *status, void *stream);
/* Attempts the last persistent-engine call executed, speculative ones
 * included (synchronises `stream`). */
int dil_sign_ctx_executed_attempts(dil_sign_ctx *ctx, uint64_t *count, void *stream);
/* Phase profile of the persistent engine (for tuning): when enabled, each
 * workgroup sums shader-clock cycles spent in its six phases (assign, y, w,
 * c, z, commit) and the sums over workgroups of the last call are returned by
 * dil_sign_ctx_phase_cycles (synchronises `stream`).  Off by default. */
int dil_sign_ctx_profile_phases(dil_sign_ctx *ctx, int enable);
/* Message hashing of the context's sign calls: 0 or DIL_HASH_LONG_MESSAGES
 * (mu and rhoprime with one sponge per 16-lane row). */
int dil_sign_ctx_set_message_hashing(dil_sign_ctx *ctx, int flags);
int dil_sign_ctx_phase_cycles(dil_sign_ctx *ctx, uint64_t *cycles6, void *stream);

/* Resident keys (no counterpart in the reference, which re-derives A from
 * rho on every signature and verification, sign.cu:98-163 and
 * verify.cu:146-200, and re-unpacks and transforms s1, s2, t0 on every
 * signature, sign.cu:306-403).  A key store holds `capacity` keys of one
 * mode: pk, sk (absent for verify-only keys), tr = H(pk) and the expanded
 * A, s1^, s2^, t0^, t1^ (dil_keystore_row_bytes(mode) each, about 38/60/95
 * KiB for modes 2/3/5).  Results equal the non-resident entry points'. */
typedef struct dil_keystore dil_keystore;
size_t dil_keystore_row_bytes(int mode);
int dil_keystore_create(dil_keystore **ks, int mode, size_t capacity);
int dil_keystore_destroy(dil_keystore *ks);
size_t dil_keystore_capacity(const dil_keystore *ks);
int dil_keystore_mode(const dil_keystore *ks);
/* 1 if key idx was loaded with its secret key or generated in the store, 0
 * for verify-only keys and keys never installed. */
int dil_keystore_has_secret(const dil_keystore *ks, size_t idx);
/* Installs n keys at [first, first + n) from pk (and sk, or NULL for
 * verify-only keys), host or device memory, and expands them on `stream`.
 * A verify-only load clears the rows' secret-key bytes and s1^, s2^, t0^. */
int dil_keystore_load(dil_keystore *ks, size_t first, const uint8_t *pk, size_t pk_pitch,
                      const uint8_t *sk, size_t sk_pitch, size_t n, void *stream);
/* Generates n keys into [first, first + n) as dil_keypair_device would
 * (explicit seeds, or counter seeds from rand_index when d_seeds is NULL)
 * and expands them. */
int dil_keystore_keygen(dil_keystore *ks, size_t first, const uint8_t *d_seeds, size_t seed_pitch,
                        uint64_t rand_index, size_t n, void *stream);
/* Copies pk and/or sk (either may be NULL) of keys [first, first + n) out;
 * DIL_EINVAL when sk is asked for and a key in the range is verify-only. */
int dil_keystore_export(const dil_keystore *ks, size_t first, size_t n, uint8_t *pk, size_t pk_pitch,
                        uint8_t *sk, size_t sk_pitch, void *stream);
/* A signing context whose rows omit the per-item expanded arrays; it only
 * accepts dil_sign_device_resident. */
int dil_sign_ctx_create_resident(dil_sign_ctx **ctx, int mode, size_t max_batch, size_t nslots);
/* dil_sign_device with item i signed by resident key d_key_idx[i] (device
 * array).  Indices must be below the store's capacity (larger ones are
 * clamped to the last key).  An item whose key is verify-only gets an
 * all-zero signature, and the call's status (dil_sign_ctx_status) has
 * DIL_SIGN_STATUS_NO_SECRET set. */
int dil_sign_device_resident(dil_sign_ctx *ctx, const dil_keystore *ks, const uint32_t *d_key_idx,
                             uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_msg,
                             const uint64_t *d_moff, const uint32_t *d_mlen, size_t n, int engine,
                             void *stream);
/* dil_verify_device with item i verified under resident key d_key_idx[i]; an
 * index at or above the capacity gives ret -1. */
int dil_verify_device_resident(const dil_keystore *ks, const uint32_t *d_key_idx, int32_t *d_ret,
                               const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_msg,
                               const uint64_t *d_moff, const uint32_t *d_mlen, size_t n,
                               uint8_t *d_work, size_t work_pitch, void *stream);
/* The same with message-hashing flags (0 or DIL_HASH_LONG_MESSAGES). */
int dil_verify_device_resident_ex(const dil_keystore *ks, const uint32_t *d_key_idx, int32_t *d_ret,
                                  const uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_msg,
                                  const uint64_t *d_moff, const uint32_t *d_mlen, size_t n,
                                  uint8_t *d_work, size_t work_pitch, int flags, void *stream);

/* Randomized signing: the reference's DILITHIUM_RANDOMIZED_SIGNING build
 * (api.cu:143-148), where rhoprime is 64 random bytes per item instead of
 * SHAKE256(key || mu).  Here the caller supplies them (d_rhoprime, one 64-byte
 * row per item) from the random source it trusts, so results are
 * reproducible from the inputs.  With ks == NULL item i signs with the
 * secret key at d_sk + i * sk_pitch (as dil_sign_device); with a key store,
 * with resident key d_key_idx[i] (d_sk unused). */
int dil_sign_device_randomized(dil_sign_ctx *ctx, const dil_keystore *ks, const uint32_t *d_key_idx,
                               uint8_t *d_sig, size_t sig_pitch, const uint8_t *d_sk, size_t sk_pitch,
                               const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                               const uint8_t *d_rhoprime, size_t rhoprime_pitch, size_t n, int engine,
                               void *stream);

/* HIP graphs of whole pipelines (no counterpart in the reference, whose
 * host scheduler synchronises once per rejection round, api.cu:208-441).
 * Each *_graph_create captures every launch of one call of the matching
 * *_device function, with these exact arguments, into an instantiated
 * graph; dil_graph_launch replays it on `stream` (one hipGraphLaunch).
 * Buffers named at creation must stay allocated while the graph exists;
 * replays of graphs sharing buffers (the same sign context, work rows or
 * outputs) must not overlap in time.  Sign graphs use the persistent
 * engine with the context's tiling at creation. */
typedef struct dil_graph dil_graph;
int dil_sign_graph_create(dil_graph **graph, dil_sign_ctx *ctx, uint8_t *d_sig, size_t sig_pitch,
                          const uint8_t *d_sk, size_t sk_pitch, const uint8_t *d_msg,
                          const uint64_t *d_moff, const uint32_t *d_mlen, size_t n);
int dil_verify_graph_create(dil_graph **graph, int mode, int32_t *d_ret, const uint8_t *d_sig,
                            size_t sig_pitch, const uint8_t *d_pk, size_t pk_pitch,
                            const uint8_t *d_msg, const uint64_t *d_moff, const uint32_t *d_mlen,
                            size_t n, uint8_t *d_work, size_t work_pitch);
int dil_keypair_graph_create(dil_graph **graph, int mode, uint8_t *d_pk, size_t pk_pitch,
                             uint8_t *d_sk, size_t sk_pitch, const uint8_t *d_seeds,
                             size_t seed_pitch, uint64_t rand_index, size_t n, uint8_t *d_work,
                             size_t work_pitch);
int dil_graph_launch(dil_graph *graph, void *stream);
size_t dil_graph_batch(const dil_graph *graph);
int dil_graph_destroy(dil_graph *graph);

/* ---------------------------------------------------------------------- */
/* Serving loop (no counterpart in the reference, whose callers batch by   */
/* hand, benches/bench_dilithium.cu:186-239).  Single requests from any    */
/* host thread are batched continuously per (op, mode) group, each group   */
/* with `lanes` batches in flight on streams of its own and device buffers */
/* preallocated from a fixed HBM budget; sign and verify use resident      */
/* keys.  Each lane's stream has a hardware queue of its own (a stream     */
/* with an all-CU mask), so lanes never serialise behind each other.       */
/* ---------------------------------------------------------------------- */

#define DIL_OP_KEYGEN 0
#define DIL_OP_SIGN 1
#define DIL_OP_VERIFY 2

typedef struct dil_server_config {
  size_t hbm_budget_bytes;  /* device bytes for in-flight batches (all lanes) */
  uint32_t lanes;           /* batches in flight per (op, mode) group (1..16) */
  uint32_t max_batch;       /* requests per batch (<= 65536) */
  uint32_t max_keys[3];     /* resident key capacity of modes 2, 3, 5 */
  uint32_t batch_window_us; /* hold a partial batch this long for more requests */
  uint32_t long_message_bytes; /* sign / verify messages above this (bytes) get
                                  lanes of their own (0: no split) */
  uint32_t long_lanes;      /* lanes per long-message (op, mode) group (0..16;
                               0: `lanes`) */
} dil_server_config;

/* One request.  The caller owns it and every buffer it points to until
 * `done` reads 1; the server writes the output fields first.
 *   keygen: seed (32 bytes) -> out = pk || sk
 *   sign:   msg, mlen, resident key `key` -> out = signature
 *   verify: msg, mlen, sig (signature bytes), resident key `key` -> verdict */
typedef struct dil_request {
  int32_t op;
  int32_t mode;
  uint32_t key;
  uint32_t mlen;
  const uint8_t *msg;
  const uint8_t *seed;
  const uint8_t *sig;
  uint8_t *out;
  int32_t status;   /* 0, or DIL_E* (DIL_ERANGE: larger than a lane admits) */
  int32_t verdict;  /* verify: 0 valid, -1 invalid */
  uint64_t t_submit_ns, t_launch_ns, t_done_ns; /* CLOCK_MONOTONIC */
  uint32_t batch_id;
  int32_t done;     /* set to 1 (release) when complete */
} dil_request;

typedef struct dil_server_stats {
  uint64_t submitted, completed, batches, rejected;
  uint64_t items_by_op[3];
  uint64_t lane_bytes;      /* device bytes of the largest lane */
  uint64_t max_msg_bytes;   /* total message bytes one sign / verify batch admits */
  /* per op (keygen, sign, verify), summed over its batches: */
  uint64_t batches_by_op[3];
  uint64_t stage_ns[3];     /* host: gather inputs into the pinned buffer, enqueue */
  uint64_t device_ns[3];    /* stream time from input copy to output copy, HIP events */
  uint64_t finish_ns[3];    /* host: scatter outputs into the requests */
} dil_server_stats;

typedef struct dil_server dil_server;
int dil_server_create(dil_server **srv, const dil_server_config *cfg);
/* Completes every submitted request, then frees the server. */
int dil_server_destroy(dil_server *srv);
/* The resident key store of `mode` (load or generate keys into it before
 * submitting requests that name them). */
dil_keystore *dil_server_keystore(dil_server *srv, int mode);
/* Queues n requests (stamps t_submit_ns). */
int dil_server_submit(dil_server *srv, dil_request *const *reqs, size_t n);
/* Waits until `req` is done (timeout_us < 0: no limit); 0 or DIL_ERANGE on
 * timeout. */
int dil_server_wait(dil_server *srv, dil_request *req, int64_t timeout_us);
/* Waits until every request submitted so far is done. */
int dil_server_drain(dil_server *srv);
int dil_server_get_stats(dil_server *srv, dil_server_stats *stats);
/* Load generator: submits reqs[i] at (start + arrival_ns[i]) from the
 * calling thread, then returns (does not wait for completion). */
int dil_server_replay(dil_server *srv, dil_request *reqs, const uint64_t *arrival_ns, size_t n);
uint64_t dil_monotonic_ns(void);

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

/* Measured lane-operations per second of one VALU instruction class on the
 * current device (synchronous): 0 add_u32, 1 bitop3_b32, 2 mul_lo_u32,
 * 3 mul_hi_i32, 4 fma_f64, 5 alignbit_b32, 6 one (cvt_f64_i32, mul_f64,
 * cvt_i32_f64, add) sequence, 7 mul_u32_u24. */
int dil_valu_microbench(int op, uint32_t iters, double *lane_ops_per_s);

/* Latency and throughput of chained Keccak-f[1600] permutations in the
 * engine's three state layouts (form 0: one state per lane, 1: two lanes per
 * state, 2: one state per 16-lane row): `blocks` workgroups of `threads`
 * (64 or 256) threads each run `nperm` dependent permutations (synchronous).
 * Returns microseconds per chained permutation and permutations per second. */
int dil_keccak_microbench(int form, uint32_t blocks, uint32_t threads, uint32_t nperm, double *us_per_perm,
                          double *perms_per_s);

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
/* A non-blocking stream with a hardware queue of its own (an all-CU mask):
 * callers that run batches on several streams at once (the reference's
 * stream mode) should create them here; streams from hipStreamCreate share
 * the runtime's few queues, where concurrent calls serialise. */
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

/* Synthetic workload bytes: the little-endian outputs of xoshiro256**
 * (Blackman-Vigna) seeded through splitmix64 from `seed`. */
int dil_xoshiro_bytes(uint64_t seed, uint8_t *out, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* CUDILITHIUM_AMD_H */
