/*
 * CPU oracle: Dilithium round-3 v3.1 (deterministic signing) restated from
 * the reference's semantics.  TEST INFRASTRUCTURE ONLY (see dil_oracle.h).
 *
 * Reference file citations are relative to the reference repository root.
 */
#include "dil_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define N 256
#define Q 8380417
#define QINV 58728449u /* q^-1 mod 2^32, reduce.cuh:12 */
#define D 13
#define SEEDBYTES 32
#define CRHBYTES 64
#define SHAKE128_RATE 168
#define SHAKE256_RATE 136
#define MAXK 8
#define MAXL 7

/* ------------------------------------------------------------------------ */
/* Parameters: include/params.h:7-73                                        */
/* ------------------------------------------------------------------------ */
int oracle_params(int mode, oracle_params_t *p) {
    memset(p, 0, sizeof(*p));
    p->mode = mode;
    switch (mode) {
        case 2:
            p->k = 4; p->l = 4; p->eta = 2; p->tau = 39; p->beta = 78;
            p->gamma1 = 1 << 17; p->gamma2 = (Q - 1) / 88; p->omega = 80;
            break;
        case 3:
            p->k = 6; p->l = 5; p->eta = 4; p->tau = 49; p->beta = 196;
            p->gamma1 = 1 << 19; p->gamma2 = (Q - 1) / 32; p->omega = 55;
            break;
        case 5:
            p->k = 8; p->l = 7; p->eta = 2; p->tau = 60; p->beta = 120;
            p->gamma1 = 1 << 19; p->gamma2 = (Q - 1) / 32; p->omega = 75;
            break;
        default:
            return -1;
    }
    p->polyeta_bytes = p->eta == 2 ? 96 : 128;
    p->polyz_bytes = p->gamma1 == (1 << 17) ? 576 : 640;
    p->polyw1_bytes = p->gamma2 == (Q - 1) / 88 ? 192 : 128;
    p->pk_bytes = SEEDBYTES + p->k * 320;
    p->sk_bytes = 3 * SEEDBYTES + (p->l + p->k) * p->polyeta_bytes + p->k * 416;
    p->sig_bytes = SEEDBYTES + p->l * p->polyz_bytes + p->omega + p->k;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Keccak-f[1600] and SHAKE, from FIPS 202 (the reference's device form is  */
/* fips202.cuh:104-273; constants fips202.cu:13-34).  The round constants   */
/* and rotation offsets are generated from the FIPS 202 definitions.        */
/* ------------------------------------------------------------------------ */
static uint64_t keccak_rc[24];
static unsigned keccak_rho[25];
static int keccak_ready = 0;

static int rc_bit(unsigned t) { /* FIPS 202 Algorithm 5 */
    unsigned r = 1;
    if (t % 255 == 0) return 1;
    for (unsigned i = 1; i <= t % 255; ++i) {
        r <<= 1;
        if (r & 0x100) r ^= 0x171; /* R[0],R[4],R[5],R[6] ^= R[8]; drop R[8] */
    }
    return r & 1;
}

static void keccak_init_tables(void) {
    if (keccak_ready) return;
    for (unsigned ir = 0; ir < 24; ++ir) {
        uint64_t rc = 0;
        for (unsigned j = 0; j < 7; ++j)
            if (rc_bit(j + 7 * ir)) rc |= 1ull << ((1u << j) - 1);
        keccak_rc[ir] = rc;
    }
    unsigned x = 1, y = 0;
    keccak_rho[0] = 0;
    for (unsigned t = 0; t < 24; ++t) {
        keccak_rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        unsigned nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx; y = ny;
    }
    keccak_ready = 1;
}

static inline uint64_t rotl64(uint64_t v, unsigned s) {
    return s ? (v << s) | (v >> (64 - s)) : v;
}

void oracle_keccakf1600(uint64_t a[25]) {
    keccak_init_tables();
    for (int round = 0; round < 24; ++round) {
        uint64_t c[5], d[5], b[25];
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
        for (int i = 0; i < 25; ++i) a[i] ^= d[i % 5];
        /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y)
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(a[x + 5 * y], keccak_rho[x + 5 * y]);
        for (int y = 0; y < 5; ++y)
            for (int x = 0; x < 5; ++x)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= keccak_rc[round];
    }
}

typedef struct {
    uint64_t s[25];
    unsigned pos;  /* byte position inside the current block */
    unsigned rate;
} sponge_t;

static void sponge_init(sponge_t *st, unsigned rate) {
    memset(st->s, 0, sizeof(st->s));
    st->pos = 0;
    st->rate = rate;
}

static void sponge_xor_byte(sponge_t *st, unsigned i, uint8_t v) {
    st->s[i / 8] ^= (uint64_t)v << (8 * (i % 8));
}

static void sponge_absorb(sponge_t *st, const uint8_t *in, size_t len) {
    for (size_t i = 0; i < len; ++i) {
        sponge_xor_byte(st, st->pos++, in[i]);
        if (st->pos == st->rate) {
            oracle_keccakf1600(st->s);
            st->pos = 0;
        }
    }
}

/* SHAKE domain byte 0x1F and final 0x80 (fips202.cuh:200-210). */
static void sponge_finalize(sponge_t *st) {
    sponge_xor_byte(st, st->pos, 0x1F);
    sponge_xor_byte(st, st->rate - 1, 0x80);
    st->pos = st->rate; /* forces a permutation before the first output byte */
}

static void sponge_squeeze(sponge_t *st, uint8_t *out, size_t len) {
    for (size_t i = 0; i < len; ++i) {
        if (st->pos == st->rate) {
            oracle_keccakf1600(st->s);
            st->pos = 0;
        }
        out[i] = (uint8_t)(st->s[st->pos / 8] >> (8 * (st->pos % 8)));
        st->pos++;
    }
}

static void shake(unsigned rate, uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    sponge_t st;
    sponge_init(&st, rate);
    sponge_absorb(&st, in, inlen);
    sponge_finalize(&st);
    sponge_squeeze(&st, out, outlen);
}

void oracle_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    shake(SHAKE128_RATE, out, outlen, in, inlen);
}

void oracle_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    shake(SHAKE256_RATE, out, outlen, in, inlen);
}

/* ------------------------------------------------------------------------ */
/* Modular arithmetic: reduce.cuh:12,42-78, ntt.cuh:64-72, rounding.cuh     */
/* ------------------------------------------------------------------------ */
int32_t oracle_montgomery_reduce(int64_t a) {
    /* t = lo32(a) * q^-1 mod 2^32 (signed); (a - t*q) is divisible by 2^32. */
    int32_t t = (int32_t)((uint32_t)(uint64_t)a * QINV);
    return (int32_t)((a - (int64_t)t * Q) >> 32);
}

int32_t oracle_reduce32(int32_t a) { /* reduce.cuh:42-48 */
    int32_t t = (a + (1 << 22)) >> 23;
    return a - t * Q;
}

int32_t oracle_caddq(int32_t a) { /* reduce.cuh:59-62 */
    return a + ((a >> 31) & Q);
}

int32_t oracle_power2round(int32_t *a0, int32_t a) { /* rounding.cuh:22-28 */
    int32_t a1 = (a + (1 << (D - 1)) - 1) >> D;
    *a0 = a - (a1 << D);
    return a1;
}

/* rounding.cuh:44-59.  a1 = round(a / (2*gamma2)) with the wrap at
 * (q-1)/(2*gamma2) mapped to 0; a0 = a - a1*2*gamma2 centred. */
int32_t oracle_decompose(int mode, int32_t *a0, int32_t a) {
    int32_t a1 = (a + 127) >> 7;
    int32_t g2;
    if (mode == 2) {
        g2 = (Q - 1) / 88;
        a1 = (a1 * 11275 + (1 << 23)) >> 24;
        a1 ^= ((43 - a1) >> 31) & a1;
    } else {
        g2 = (Q - 1) / 32;
        a1 = (a1 * 1025 + (1 << 21)) >> 22;
        a1 &= 15;
    }
    *a0 = a - a1 * 2 * g2;
    *a0 -= (((Q - 1) / 2 - *a0) >> 31) & Q;
    return a1;
}

int32_t oracle_make_hint(int mode, int32_t a0, int32_t a1) { /* rounding.cuh:70-75 */
    int32_t g2 = mode == 2 ? (Q - 1) / 88 : (Q - 1) / 32;
    return (a0 > g2 || a0 < -g2 || (a0 == -g2 && a1 != 0)) ? 1 : 0;
}

int32_t oracle_use_hint(int mode, int32_t a, int32_t hint) { /* rounding.cuh:86-107 */
    int32_t a0, a1 = oracle_decompose(mode, &a0, a);
    if (hint == 0) return a1;
    if (mode == 2) {
        if (a0 > 0) return a1 == 43 ? 0 : a1 + 1;
        return a1 == 0 ? 43 : a1 - 1;
    }
    return a0 > 0 ? (a1 + 1) & 15 : (a1 - 1) & 15;
}

void oracle_elementwise(int mode, int op, const int32_t *a, const int32_t *b, int32_t *r0,
                        int32_t *r1, size_t n) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        int32_t lo = 0, hi = 0;
        switch (op) {
            case 0: hi = oracle_power2round(&lo, a[i]); break;
            case 1: hi = oracle_decompose(mode, &lo, a[i]); break;
            case 2: hi = oracle_make_hint(mode, a[i], b[i]); break;
            case 3: hi = oracle_use_hint(mode, a[i], b[i]); break;
            case 4: hi = oracle_reduce32(a[i]); break;
            case 5: hi = oracle_caddq(a[i]); break;
            default: hi = oracle_montgomery_reduce((int64_t)a[i] * b[i]); break;
        }
        r1[i] = hi;
        if (r0) r0[i] = lo;
    }
}

/* ------------------------------------------------------------------------ */
/* NTT (ntt.cu:24-207; same butterfly DAG as the radix-8 form)              */
/* zetas[i] = centred(MONT * 1753^brv8(i) mod q), zetas[0] = 0              */
/* ------------------------------------------------------------------------ */
static int32_t zetas_tab[N];
static int zetas_ready = 0;

static void zetas_init(void) {
    if (zetas_ready) return;
    const uint64_t mont = ((uint64_t)1 << 32) % Q;
    zetas_tab[0] = 0;
    for (unsigned i = 1; i < N; ++i) {
        unsigned br = 0;
        for (unsigned b = 0; b < 8; ++b) br |= ((i >> b) & 1u) << (7 - b);
        uint64_t p = 1, base = 1753;
        for (unsigned e = br; e; e >>= 1) {
            if (e & 1) p = p * base % Q;
            base = base * base % Q;
        }
        int64_t v = (int64_t)(p * mont % Q);
        if (v > (Q - 1) / 2) v -= Q;
        zetas_tab[i] = (int32_t)v;
    }
    zetas_ready = 1;
}

const int32_t *oracle_zetas(void) {
    zetas_init();
    return zetas_tab;
}

/* Forward: Cooley-Tukey, len 128..1, zeta index k = 1..255 in block order;
 * butterfly (a, b) -> (a + mont(z*b), a - mont(z*b)); no intermediate
 * reduction (ntt.cu:58-63, 77-140). */
void oracle_ntt(int32_t a[N]) {
    zetas_init();
    unsigned k = 0;
    for (unsigned len = 128; len > 0; len >>= 1) {
        for (unsigned start = 0; start < N; start += 2 * len) {
            int32_t z = zetas_tab[++k];
            for (unsigned j = start; j < start + len; ++j) {
                int32_t t = oracle_montgomery_reduce((int64_t)z * a[j + len]);
                a[j + len] = a[j] - t;
                a[j] = a[j] + t;
            }
        }
    }
}

/* Inverse: Gentleman-Sande with -zetas in reverse order, then scaling by
 * f = MONT^2/256 = 41978 (ntt.cu:65-69, 142-207). */
void oracle_invntt_tomont(int32_t a[N]) {
    zetas_init();
    const int32_t f = 41978;
    unsigned k = N;
    for (unsigned len = 1; len < N; len <<= 1) {
        for (unsigned start = 0; start < N; start += 2 * len) {
            int32_t z = -zetas_tab[--k];
            for (unsigned j = start; j < start + len; ++j) {
                int32_t t = a[j];
                a[j] = t + a[j + len];
                a[j + len] = oracle_montgomery_reduce((int64_t)z * (t - a[j + len]));
            }
        }
    }
    for (unsigned j = 0; j < N; ++j) a[j] = oracle_montgomery_reduce((int64_t)f * a[j]);
}

void oracle_ntt_pipeline(int32_t *c, const int32_t *a, const int32_t *b, size_t npoly) {
    zetas_init();
#pragma omp parallel for schedule(static)
    for (long long p = 0; p < (long long)npoly; ++p) {
        int32_t t[N];
        memcpy(t, a + (size_t)p * N, sizeof(t));
        oracle_ntt(t);
        for (int i = 0; i < N; ++i)
            t[i] = oracle_montgomery_reduce((int64_t)t[i] * b[(size_t)p * N + i]);
        oracle_invntt_tomont(t);
        memcpy(c + (size_t)p * N, t, sizeof(t));
    }
}

void oracle_ntt_batch(int32_t *a, size_t npoly, int inverse) {
    zetas_init();
#pragma omp parallel for schedule(static)
    for (long long p = 0; p < (long long)npoly; ++p) {
        if (inverse)
            oracle_invntt_tomont(a + (size_t)p * N);
        else
            oracle_ntt(a + (size_t)p * N);
    }
}

/* ------------------------------------------------------------------------ */
/* Little-endian bit packing.  Every byte format on the path (t1 10-bit,    */
/* t0 13-bit, eta 3/4-bit, z 18/20-bit, w1 6/4-bit) is the plain LE bit     */
/* stream of 256 unsigned fields (keypair.cu:116-130,268-301; sign.cu:      */
/* 662-683,1151-1192).                                                      */
/* ------------------------------------------------------------------------ */
static void pack_bits(uint8_t *out, const uint32_t *v, unsigned w) {
    memset(out, 0, N * w / 8);
    for (unsigned i = 0; i < N; ++i)
        for (unsigned b = 0; b < w; ++b)
            if ((v[i] >> b) & 1u) {
                unsigned pos = i * w + b;
                out[pos / 8] |= (uint8_t)(1u << (pos % 8));
            }
}

static void unpack_bits(uint32_t *v, const uint8_t *in, unsigned w) {
    for (unsigned i = 0; i < N; ++i) {
        uint32_t x = 0;
        for (unsigned b = 0; b < w; ++b) {
            unsigned pos = i * w + b;
            x |= (uint32_t)((in[pos / 8] >> (pos % 8)) & 1u) << b;
        }
        v[i] = x;
    }
}

/* offset - a, packed in w bits */
static void pack_offset(uint8_t *out, const int32_t *a, int32_t offset, unsigned w) {
    uint32_t v[N];
    for (int i = 0; i < N; ++i) v[i] = (uint32_t)(offset - a[i]);
    pack_bits(out, v, w);
}

static void unpack_offset(int32_t *a, const uint8_t *in, int32_t offset, unsigned w) {
    uint32_t v[N];
    unpack_bits(v, in, w);
    for (int i = 0; i < N; ++i) a[i] = offset - (int32_t)v[i];
}

/* ------------------------------------------------------------------------ */
/* Samplers                                                                  */
/* ------------------------------------------------------------------------ */

/* ExpandA entry: rejection-uniform over SHAKE128(rho || le16(nonce)), 23-bit
 * LE candidates, accept < q (sign.cu:98-163; keypair.cu:140-189).  The
 * stream continues past 5 blocks as the spec does; the reference reads 24
 * bytes beyond its 840-byte buffer instead (probability ~0). */
static void sample_uniform(int32_t *a, const uint8_t rho[SEEDBYTES], uint16_t nonce) {
    uint8_t in[SEEDBYTES + 2];
    memcpy(in, rho, SEEDBYTES);
    in[SEEDBYTES] = (uint8_t)nonce;
    in[SEEDBYTES + 1] = (uint8_t)(nonce >> 8);
    sponge_t st;
    sponge_init(&st, SHAKE128_RATE);
    sponge_absorb(&st, in, sizeof(in));
    sponge_finalize(&st);
    int ctr = 0;
    while (ctr < N) {
        uint8_t b[3];
        sponge_squeeze(&st, b, 3);
        uint32_t t = b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)(b[2] & 0x7F) << 16);
        if (t < Q) a[ctr++] = (int32_t)t;
    }
}

/* ExpandS entry: SHAKE256(rhoprime || le16(nonce)), nibbles low-first,
 * eta=2: accept < 15, 2 - (t mod 5); eta=4: accept < 9, 4 - t
 * (keypair.cu:80-115). */
static void sample_eta(int32_t *a, const uint8_t rhoprime[CRHBYTES], uint16_t nonce, int eta) {
    uint8_t in[CRHBYTES + 2];
    memcpy(in, rhoprime, CRHBYTES);
    in[CRHBYTES] = (uint8_t)nonce;
    in[CRHBYTES + 1] = (uint8_t)(nonce >> 8);
    sponge_t st;
    sponge_init(&st, SHAKE256_RATE);
    sponge_absorb(&st, in, sizeof(in));
    sponge_finalize(&st);
    int ctr = 0;
    while (ctr < N) {
        uint8_t byte;
        sponge_squeeze(&st, &byte, 1);
        uint32_t t[2] = {byte & 0x0Fu, (uint32_t)byte >> 4};
        for (int h = 0; h < 2 && ctr < N; ++h) {
            if (eta == 2 && t[h] < 15)
                a[ctr++] = 2 - (int32_t)(t[h] % 5);
            else if (eta == 4 && t[h] < 9)
                a[ctr++] = 4 - (int32_t)t[h];
        }
    }
}

/* ExpandMask entry: SHAKE256(rhoprime || le16(nonce)), 5 blocks, unpacked as
 * gamma1 - (18- or 20-bit LE field) (sign.cu:463-526). */
static void sample_gamma1(int32_t *a, const uint8_t rhoprime[CRHBYTES], uint16_t nonce,
                          int gamma1) {
    uint8_t in[CRHBYTES + 2], buf[5 * SHAKE256_RATE];
    memcpy(in, rhoprime, CRHBYTES);
    in[CRHBYTES] = (uint8_t)nonce;
    in[CRHBYTES + 1] = (uint8_t)(nonce >> 8);
    shake(SHAKE256_RATE, buf, sizeof(buf), in, sizeof(in));
    unpack_offset(a, buf, gamma1, gamma1 == (1 << 17) ? 18 : 20);
}

/* SampleInBall (sign.cu:710-734; verify.cu:221-245).  The spec squeezes a
 * further block when the 136-byte block runs out; the reference does not
 * (probability ~0 for tau <= 60). */
static void sample_challenge(int32_t c[N], const uint8_t ctilde[SEEDBYTES], int tau) {
    sponge_t st;
    uint8_t buf[8];
    sponge_init(&st, SHAKE256_RATE);
    sponge_absorb(&st, ctilde, SEEDBYTES);
    sponge_finalize(&st);
    sponge_squeeze(&st, buf, 8);
    uint64_t signs = 0;
    for (int i = 0; i < 8; ++i) signs |= (uint64_t)buf[i] << (8 * i);
    memset(c, 0, sizeof(int32_t) * N);
    for (int i = N - tau; i < N; ++i) {
        uint8_t b;
        do {
            sponge_squeeze(&st, &b, 1);
        } while (b > i);
        c[i] = c[b];
        c[b] = 1 - 2 * (int32_t)(signs & 1);
        signs >>= 1;
    }
}

static int chknorm(const int32_t *a, int32_t bound) { /* poly.cuh:31-54 */
    if (bound > (Q - 1) / 8) return 1;
    for (int i = 0; i < N; ++i) {
        int32_t t = a[i] >> 31;
        t = a[i] - (t & 2 * a[i]);
        if (t >= bound) return 1;
    }
    return 0;
}

static void poly_pointwise(int32_t *c, const int32_t *a, const int32_t *b) {
    for (int i = 0; i < N; ++i) c[i] = oracle_montgomery_reduce((int64_t)a[i] * b[i]);
}

static void expand_matrix(int32_t *mat, const uint8_t rho[SEEDBYTES], int k, int l) {
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < l; ++j)
            sample_uniform(mat + (size_t)(i * l + j) * N, rho, (uint16_t)((i << 8) + j));
}

/* acc = sum_j mont(mat[i][j] * v[j]) for row i */
static void matrix_row_acc(int32_t *acc, const int32_t *mat, const int32_t *v, int i, int l) {
    int32_t t[N];
    memset(acc, 0, sizeof(int32_t) * N);
    for (int j = 0; j < l; ++j) {
        poly_pointwise(t, mat + (size_t)(i * l + j) * N, v + (size_t)j * N);
        for (int x = 0; x < N; ++x) acc[x] += t[x];
    }
}

/* ------------------------------------------------------------------------ */
/* Keygen: keypair.cu:35-311, api.cu:26-61                                  */
/* ------------------------------------------------------------------------ */
void oracle_keypair_from_seed(int mode, uint8_t *pk, uint8_t *sk, const uint8_t seed[32]) {
    oracle_params_t p;
    oracle_params(mode, &p);
    uint8_t seedbuf[2 * SEEDBYTES + CRHBYTES];
    /* rho || rhoprime || key = SHAKE256(seed) (keypair.cu:62-67) */
    oracle_shake256(seedbuf, sizeof(seedbuf), seed, SEEDBYTES);
    const uint8_t *rho = seedbuf, *rhoprime = seedbuf + SEEDBYTES, *key = rhoprime + CRHBYTES;

    int32_t s1[MAXL * N], s1hat[MAXL * N], s2[MAXK * N], mat[MAXK * MAXL * N];
    for (int j = 0; j < p.l; ++j) sample_eta(s1 + j * N, rhoprime, (uint16_t)j, p.eta);
    for (int i = 0; i < p.k; ++i) sample_eta(s2 + i * N, rhoprime, (uint16_t)(p.l + i), p.eta);
    expand_matrix(mat, rho, p.k, p.l);
    memcpy(s1hat, s1, sizeof(int32_t) * p.l * N);
    for (int j = 0; j < p.l; ++j) oracle_ntt(s1hat + j * N);

    memcpy(pk, rho, SEEDBYTES);
    uint8_t *sk_s1 = sk + 3 * SEEDBYTES;
    uint8_t *sk_s2 = sk_s1 + p.l * p.polyeta_bytes;
    uint8_t *sk_t0 = sk_s2 + p.k * p.polyeta_bytes;
    const unsigned wbits = p.eta == 2 ? 3 : 4;
    for (int j = 0; j < p.l; ++j) pack_offset(sk_s1 + j * p.polyeta_bytes, s1 + j * N, p.eta, wbits);
    for (int i = 0; i < p.k; ++i) pack_offset(sk_s2 + i * p.polyeta_bytes, s2 + i * N, p.eta, wbits);

    for (int i = 0; i < p.k; ++i) {
        int32_t t[N], t1[N], t0[N];
        matrix_row_acc(t, mat, s1hat, i, p.l);
        for (int x = 0; x < N; ++x) t[x] = oracle_reduce32(t[x]);
        oracle_invntt_tomont(t);
        /* + s2, caddq, Power2Round (keypair.cu:251-264) */
        for (int x = 0; x < N; ++x) {
            int32_t v = oracle_caddq(t[x] + s2[i * N + x]);
            t1[x] = oracle_power2round(&t0[x], v);
        }
        uint32_t u[N];
        for (int x = 0; x < N; ++x) u[x] = (uint32_t)t1[x];
        pack_bits(pk + SEEDBYTES + i * 320, u, 10);
        pack_offset(sk_t0 + i * 416, t0, 1 << (D - 1), 13);
    }
    /* sk = rho || key || tr, tr = SHAKE256(pk)[0:32] (keypair.cu:305-310) */
    memcpy(sk, rho, SEEDBYTES);
    memcpy(sk + SEEDBYTES, key, SEEDBYTES);
    oracle_shake256(sk + 2 * SEEDBYTES, SEEDBYTES, pk, (size_t)p.pk_bytes);
}

void oracle_keypair_counter(int mode, uint8_t *pk, uint8_t *sk, uint64_t ctr) {
    uint8_t c[8], seed[SEEDBYTES];
    for (int i = 0; i < 8; ++i) c[i] = (uint8_t)(ctr >> (8 * i));
    oracle_shake128(seed, SEEDBYTES, c, 8);
    oracle_keypair_from_seed(mode, pk, sk, seed);
}

/* ------------------------------------------------------------------------ */
/* Sign: api.cu:63-452 (fixed stage + host rejection scheduler) and the     */
/* kernels of sign.cu.  Result = attempt with the first successful nonce.   */
/* ------------------------------------------------------------------------ */
int oracle_sign(int mode, uint8_t *sig, const uint8_t *m, size_t mlen, const uint8_t *sk) {
    return oracle_sign_rnd(mode, sig, m, mlen, sk, NULL);
}

/* rnd != NULL: randomized signing, rhoprime = rnd (64 bytes), as the
 * reference does under DILITHIUM_RANDOMIZED_SIGNING (api.cu:143-148, there
 * from randombytes); NULL: deterministic rhoprime = SHAKE256(key || mu). */
int oracle_sign_rnd(int mode, uint8_t *sig, const uint8_t *m, size_t mlen, const uint8_t *sk,
                    const uint8_t *rnd) {
    oracle_params_t p;
    oracle_params(mode, &p);
    const uint8_t *rho = sk, *key = sk + SEEDBYTES, *tr = sk + 2 * SEEDBYTES;
    const uint8_t *sk_s1 = sk + 3 * SEEDBYTES;
    const uint8_t *sk_s2 = sk_s1 + p.l * p.polyeta_bytes;
    const uint8_t *sk_t0 = sk_s2 + p.k * p.polyeta_bytes;
    const unsigned wbits = p.eta == 2 ? 3 : 4;

    int32_t s1[MAXL * N], s2[MAXK * N], t0[MAXK * N], mat[MAXK * MAXL * N];
    for (int j = 0; j < p.l; ++j) {
        unpack_offset(s1 + j * N, sk_s1 + j * p.polyeta_bytes, p.eta, wbits);
        oracle_ntt(s1 + j * N);
    }
    for (int i = 0; i < p.k; ++i) {
        unpack_offset(s2 + i * N, sk_s2 + i * p.polyeta_bytes, p.eta, wbits);
        oracle_ntt(s2 + i * N);
        unpack_offset(t0 + i * N, sk_t0 + i * 416, 1 << (D - 1), 13);
        oracle_ntt(t0 + i * N);
    }
    expand_matrix(mat, rho, p.k, p.l);

    /* mu = SHAKE256(tr || m) (api.cu:137), rhoprime = SHAKE256(key || mu)
     * (api.cu:150-152) */
    uint8_t keymu[SEEDBYTES + CRHBYTES], rhoprime[CRHBYTES];
    uint8_t *mu = keymu + SEEDBYTES;
    {
        sponge_t st;
        sponge_init(&st, SHAKE256_RATE);
        sponge_absorb(&st, tr, SEEDBYTES);
        sponge_absorb(&st, m, mlen);
        sponge_finalize(&st);
        sponge_squeeze(&st, mu, CRHBYTES);
    }
    memcpy(keymu, key, SEEDBYTES);
    if (rnd)
        memcpy(rhoprime, rnd, CRHBYTES);
    else
        oracle_shake256(rhoprime, CRHBYTES, keymu, sizeof(keymu));

    const unsigned w1bits = p.gamma2 == (Q - 1) / 88 ? 6 : 4;
    const unsigned zbits = p.gamma1 == (1 << 17) ? 18 : 20;
    for (unsigned nonce = 0;; ++nonce) {
        int32_t y[MAXL * N], yhat[MAXL * N], w1[MAXK * N], w0[MAXK * N];
        int32_t c[N], chat[N], z[MAXL * N];
        uint8_t hashin[CRHBYTES + MAXK * 192];
        /* y = ExpandMask(rhoprime, L*nonce + l) with u16 nonce (sign.cu:488-498) */
        for (int j = 0; j < p.l; ++j)
            sample_gamma1(y + j * N, rhoprime, (uint16_t)(p.l * nonce + j), p.gamma1);
        memcpy(yhat, y, sizeof(int32_t) * p.l * N);
        for (int j = 0; j < p.l; ++j) oracle_ntt(yhat + j * N);
        /* w = INTT(A yhat), caddq, Decompose, pack w1 (sign.cu:605-685) */
        memcpy(hashin, mu, CRHBYTES);
        for (int i = 0; i < p.k; ++i) {
            int32_t w[N];
            matrix_row_acc(w, mat, yhat, i, p.l);
            for (int x = 0; x < N; ++x) w[x] = oracle_reduce32(w[x]);
            oracle_invntt_tomont(w);
            uint32_t u[N];
            for (int x = 0; x < N; ++x) {
                w1[i * N + x] = oracle_decompose(mode, &w0[i * N + x], oracle_caddq(w[x]));
                u[x] = (uint32_t)w1[i * N + x];
            }
            pack_bits(hashin + CRHBYTES + i * p.polyw1_bytes, u, w1bits);
        }
        /* c~ = SHAKE256(mu || w1packed)[0:32]; c = SampleInBall(c~)
         * (sign.cu:687-735) */
        oracle_shake256(sig, SEEDBYTES, hashin, CRHBYTES + (size_t)p.k * p.polyw1_bytes);
        sample_challenge(c, sig, p.tau);
        memcpy(chat, c, sizeof(chat));
        oracle_ntt(chat);

        /* rejection checks, MakeHint (sign.cu:986-1196) */
        int reject = 0;
        for (int j = 0; j < p.l && !reject; ++j) {
            int32_t *zj = z + j * N;
            poly_pointwise(zj, chat, s1 + j * N);
            oracle_invntt_tomont(zj);
            for (int x = 0; x < N; ++x) zj[x] = oracle_reduce32(zj[x] + y[j * N + x]);
            reject = chknorm(zj, p.gamma1 - p.beta);
        }
        if (reject) continue;

        uint8_t hint[MAXK + 80];
        memset(hint, 0, sizeof(hint));
        int pos = 0;
        for (int i = 0; i < p.k && !reject; ++i) {
            int32_t r0[N], ct0[N];
            poly_pointwise(r0, chat, s2 + i * N);
            oracle_invntt_tomont(r0);
            for (int x = 0; x < N; ++x) r0[x] = oracle_reduce32(w0[i * N + x] - r0[x]);
            if (chknorm(r0, p.gamma2 - p.beta)) { reject = 1; break; }
            poly_pointwise(ct0, chat, t0 + i * N);
            oracle_invntt_tomont(ct0);
            for (int x = 0; x < N; ++x) ct0[x] = oracle_reduce32(ct0[x]);
            if (chknorm(ct0, p.gamma2)) { reject = 1; break; }
            for (int x = 0; x < N; ++x) {
                if (oracle_make_hint(mode, r0[x] + ct0[x], w1[i * N + x])) {
                    if (pos >= p.omega) { reject = 1; break; }
                    hint[pos++] = (uint8_t)x;
                }
            }
            hint[p.omega + i] = (uint8_t)pos;
        }
        if (reject) continue;

        /* sig = c~ || z (gamma1 - z, 18/20 bits) || h */
        for (int j = 0; j < p.l; ++j)
            pack_offset(sig + SEEDBYTES + j * p.polyz_bytes, z + j * N, p.gamma1, zbits);
        memcpy(sig + SEEDBYTES + p.l * p.polyz_bytes, hint, (size_t)(p.omega + p.k));
        return (int)nonce + 1;
    }
}

/* ------------------------------------------------------------------------ */
/* Verify: verify.cu:28-325, api.cu:454-515                                 */
/* ------------------------------------------------------------------------ */
int oracle_verify(int mode, const uint8_t *sig, size_t siglen, const uint8_t *m, size_t mlen,
                  const uint8_t *pk) {
    oracle_params_t p;
    oracle_params(mode, &p);
    if (siglen != (size_t)p.sig_bytes) return -1; /* api.cu:461-464 */

    /* hint decode (verify.cu:52-93); spec semantics: any malformed entry
     * rejects */
    const uint8_t *h = sig + SEEDBYTES + p.l * p.polyz_bytes;
    int32_t hint[MAXK * N];
    memset(hint, 0, sizeof(hint));
    int k = 0;
    for (int i = 0; i < p.k; ++i) {
        int end = h[p.omega + i];
        if (end < k || end > p.omega) return -1;
        for (int j = k; j < end; ++j) {
            if (j > k && h[j] <= h[j - 1]) return -1;
            hint[i * N + h[j]] = 1;
        }
        k = end;
    }
    for (int j = k; j < p.omega; ++j)
        if (h[j]) return -1;

    const unsigned zbits = p.gamma1 == (1 << 17) ? 18 : 20;
    int32_t z[MAXL * N], mat[MAXK * MAXL * N], c[N];
    for (int j = 0; j < p.l; ++j) {
        unpack_offset(z + j * N, sig + SEEDBYTES + j * p.polyz_bytes, p.gamma1, zbits);
        if (chknorm(z + j * N, p.gamma1 - p.beta)) return -1;
        oracle_ntt(z + j * N);
    }

    /* tr = SHAKE256(pk)[0:32], mu = SHAKE256(tr || m) (verify.cu:130-144) */
    uint8_t tr[SEEDBYTES], hashin[CRHBYTES + MAXK * 192], ctilde[SEEDBYTES];
    oracle_shake256(tr, SEEDBYTES, pk, (size_t)p.pk_bytes);
    {
        sponge_t st;
        sponge_init(&st, SHAKE256_RATE);
        sponge_absorb(&st, tr, SEEDBYTES);
        sponge_absorb(&st, m, mlen);
        sponge_finalize(&st);
        sponge_squeeze(&st, hashin, CRHBYTES);
    }
    expand_matrix(mat, pk, p.k, p.l);
    sample_challenge(c, sig, p.tau);
    oracle_ntt(c);

    const unsigned w1bits = p.gamma2 == (Q - 1) / 88 ? 6 : 4;
    for (int i = 0; i < p.k; ++i) {
        int32_t w[N], t1[N];
        uint32_t u[N];
        matrix_row_acc(w, mat, z, i, p.l);
        unpack_bits(u, pk + SEEDBYTES + i * 320, 10);
        for (int x = 0; x < N; ++x) t1[x] = (int32_t)(u[x] << D);
        oracle_ntt(t1);
        poly_pointwise(t1, c, t1);
        for (int x = 0; x < N; ++x) w[x] = oracle_reduce32(w[x] - t1[x]);
        oracle_invntt_tomont(w);
        for (int x = 0; x < N; ++x)
            u[x] = (uint32_t)oracle_use_hint(mode, oracle_caddq(w[x]), hint[i * N + x]);
        pack_bits(hashin + CRHBYTES + i * p.polyw1_bytes, u, w1bits);
    }
    oracle_shake256(ctilde, SEEDBYTES, hashin, CRHBYTES + (size_t)p.k * p.polyw1_bytes);
    return memcmp(ctilde, sig, SEEDBYTES) == 0 ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* Batched forms (the CPU baseline: OpenMP across host cores)               */
/* ------------------------------------------------------------------------ */
static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_keypair_batch(int mode, uint8_t *pk, uint8_t *sk, const uint8_t *seeds, size_t n,
                          int nthreads) {
    oracle_params_t p;
    oracle_params(mode, &p);
    keccak_init_tables();
    zetas_init();
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (long long i = 0; i < (long long)n; ++i)
        oracle_keypair_from_seed(mode, pk + (size_t)i * p.pk_bytes, sk + (size_t)i * p.sk_bytes,
                                 seeds + (size_t)i * 32);
}

void oracle_keypair_counter_batch(int mode, uint8_t *pk, uint8_t *sk, uint64_t rand_index,
                                  size_t n, int nthreads) {
    oracle_params_t p;
    oracle_params(mode, &p);
    keccak_init_tables();
    zetas_init();
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (long long i = 0; i < (long long)n; ++i)
        oracle_keypair_counter(mode, pk + (size_t)i * p.pk_bytes, sk + (size_t)i * p.sk_bytes,
                               rand_index + 2 * (uint64_t)i + 1);
}

void oracle_sign_batch(int mode, uint8_t *sig, const uint8_t *msgs, const uint64_t *moff,
                       const uint64_t *mlen, const uint8_t *sk, size_t n, int32_t *attempts,
                       int nthreads) {
    oracle_params_t p;
    oracle_params(mode, &p);
    keccak_init_tables();
    zetas_init();
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (long long i = 0; i < (long long)n; ++i) {
        int a = oracle_sign(mode, sig + (size_t)i * p.sig_bytes, msgs + moff[i], (size_t)mlen[i],
                            sk + (size_t)i * p.sk_bytes);
        if (attempts) attempts[i] = a;
    }
}

void oracle_sign_rnd_batch(int mode, uint8_t *sig, const uint8_t *msgs, const uint64_t *moff,
                           const uint64_t *mlen, const uint8_t *sk, const uint8_t *rnd, size_t n,
                           int32_t *attempts, int nthreads) {
    oracle_params_t p;
    oracle_params(mode, &p);
    keccak_init_tables();
    zetas_init();
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (long long i = 0; i < (long long)n; ++i) {
        int a = oracle_sign_rnd(mode, sig + (size_t)i * p.sig_bytes, msgs + moff[i], (size_t)mlen[i],
                                sk + (size_t)i * p.sk_bytes, rnd + (size_t)i * CRHBYTES);
        if (attempts) attempts[i] = a;
    }
}

void oracle_verify_batch(int mode, int32_t *ret, const uint8_t *sig, const uint8_t *msgs,
                         const uint64_t *moff, const uint64_t *mlen, const uint8_t *pk, size_t n,
                         int nthreads) {
    oracle_params_t p;
    oracle_params(mode, &p);
    keccak_init_tables();
    zetas_init();
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (long long i = 0; i < (long long)n; ++i)
        ret[i] = oracle_verify(mode, sig + (size_t)i * p.sig_bytes, (size_t)p.sig_bytes,
                               msgs + moff[i], (size_t)mlen[i], pk + (size_t)i * p.pk_bytes);
}
