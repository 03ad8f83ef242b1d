/*
 * CPU oracle for the batched Dilithium hot path (round-3 v3.1 parameterisation).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU engine is compared
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links or calls it.
 *
 * It restates the semantics of the reference (encryptorion-lab/cuDilithium)
 * in portable C: every function cites the reference file:line it follows.
 * Pins (see tests/test_oracle_*.py):
 *   - whole path (keygen/sign/verify, modes 2/3/5): the SHA-256 pins of the
 *     reference's tvecs files, tests/CMakeLists.txt:21-23, reproduced by
 *     regenerating the 10,000-vector files in the format of
 *     tests/test_vectors.cu:177-209;
 *   - SHAKE128/256: Python hashlib;
 *   - zetas: closed form plus entries of src/ntt.cu:24-56;
 *   - NTT/INTT: O(N^2) evaluation at the odd powers of 1753 and the
 *     INTT(NTT(a)) == MONT*a invariant;
 *   - rounding/reduction: exhaustive checks over [0, q) against their
 *     defining equations.
 */
#ifndef DIL_ORACLE_H
#define DIL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int mode, k, l, eta, tau, beta, gamma1, gamma2, omega;
    int pk_bytes, sk_bytes, sig_bytes;
    int polyeta_bytes, polyz_bytes, polyw1_bytes;
} oracle_params_t;

/* Fills *p for mode 2, 3 or 5; returns 0, or -1 for an unknown mode. */
int oracle_params(int mode, oracle_params_t *p);

/* FIPS 202 SHAKE (one-shot). */
void oracle_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void oracle_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void oracle_keccakf1600(uint64_t st[25]);

/* Arithmetic mod q = 8380417 (reduce.cuh, rounding.cuh, ntt.cuh:64-72). */
int32_t oracle_montgomery_reduce(int64_t a);
int32_t oracle_reduce32(int32_t a);
int32_t oracle_caddq(int32_t a);
int32_t oracle_power2round(int32_t *a0, int32_t a);
int32_t oracle_decompose(int mode, int32_t *a0, int32_t a);
int32_t oracle_make_hint(int mode, int32_t a0, int32_t a1);
int32_t oracle_use_hint(int mode, int32_t a, int32_t hint);
const int32_t *oracle_zetas(void);

/* Elementwise forms over arrays (used by the exhaustive tests).
 * op: 0 power2round, 1 decompose, 2 make_hint(a, b), 3 use_hint(a, b),
 *     4 reduce32, 5 caddq, 6 montgomery_reduce((int64)a * b). */
void oracle_elementwise(int mode, int op, const int32_t *a, const int32_t *b, int32_t *r0,
                        int32_t *r1, size_t n);

/* In-place NTT / inverse NTT with Montgomery factor (ntt.cu:77-207). */
void oracle_ntt(int32_t a[256]);
void oracle_invntt_tomont(int32_t a[256]);

/* Standalone NTT workload: c = INTT(NTT(a) o_mont b), per polynomial,
 * over npoly polynomials, OpenMP-parallel.  int32 bit-exact. */
void oracle_ntt_pipeline(int32_t *c, const int32_t *a, const int32_t *b, size_t npoly);
void oracle_ntt_batch(int32_t *a, size_t npoly, int inverse);

/* Keygen from an explicit 32-byte seed (keypair.cu:62-311). */
void oracle_keypair_from_seed(int mode, uint8_t *pk, uint8_t *sk, const uint8_t seed[32]);
/* Keygen whose seed is SHAKE128(le64(ctr))[0:32] (keypair.cu:50-60). */
void oracle_keypair_counter(int mode, uint8_t *pk, uint8_t *sk, uint64_t ctr);
/* Deterministic signature (api.cu:63-452 + sign.cu).  Returns the number
 * of attempts (the first successful nonce + 1). */
int oracle_sign(int mode, uint8_t *sig, const uint8_t *m, size_t mlen, const uint8_t *sk);
/* Randomized signature with the given 64-byte rhoprime (api.cu:143-148);
 * rnd == NULL gives oracle_sign's deterministic signature. */
int oracle_sign_rnd(int mode, uint8_t *sig, const uint8_t *m, size_t mlen, const uint8_t *sk,
                    const uint8_t *rnd);
/* 0 if valid, -1 otherwise (api.cu:454-515, verify.cu:28-325). */
int oracle_verify(int mode, const uint8_t *sig, size_t siglen, const uint8_t *m, size_t mlen,
                  const uint8_t *pk);

/* Batched, OpenMP-parallel forms (schedule(dynamic)).  Messages are given as
 * one byte buffer plus per-item offsets and lengths.  nthreads <= 0 means the
 * OpenMP default. */
void oracle_keypair_batch(int mode, uint8_t *pk, uint8_t *sk, const uint8_t *seeds, size_t n,
                          int nthreads);
void oracle_keypair_counter_batch(int mode, uint8_t *pk, uint8_t *sk, uint64_t rand_index,
                                  size_t n, int nthreads);
void oracle_sign_batch(int mode, uint8_t *sig, const uint8_t *msgs, const uint64_t *moff,
                       const uint64_t *mlen, const uint8_t *sk, size_t n, int32_t *attempts,
                       int nthreads);
void oracle_sign_rnd_batch(int mode, uint8_t *sig, const uint8_t *msgs, const uint64_t *moff,
                           const uint64_t *mlen, const uint8_t *sk, const uint8_t *rnd, size_t n,
                           int32_t *attempts, int nthreads);
void oracle_verify_batch(int mode, int32_t *ret, const uint8_t *sig, const uint8_t *msgs,
                         const uint64_t *moff, const uint64_t *mlen, const uint8_t *pk, size_t n,
                         int nthreads);

/* Number of OpenMP threads the batch functions use by default. */
int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
