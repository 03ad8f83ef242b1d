// Device-side stages of deterministic signing, shared by the per-stage
// (one launch per stage, host-driven rejection rounds) form and the
// persistent work-queue engine.
//
// Fixed part per item (api.cu:117-182):
//   mu = SHAKE256(tr || m), rhoprime = SHAKE256(key || mu)   lane per item
//   A = ExpandA(rho)                                         lane per entry
//   s1^, s2^, t0^ = NTT(unpack(sk))                          wave per poly
// One attempt (item, nonce) in a "slot" (sign.cu:463-1196):
//   y    lane per (slot, l)   raw ExpandMask squeeze of rhoprime || L*nonce+l
//   w    wave per slot        NTT(y), A y^, INTT, caddq -> w; pack w1
//   c    lane per slot        c~ = SHAKE256(mu || w1), c = SampleInBall(c~)
//   z    wave per slot        z = y + INTT(c^ s1^) and the three norm checks,
//                             MakeHint, pack z and h
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"
#include "pipelines.hpp"
#include "poly.hpp"
#include "sample.hpp"

namespace dil {

struct SignArgs {
    uint8_t *sig;  // final signatures
    size_t sig_pitch;
    const uint8_t *sk;
    size_t sk_pitch;
    MsgRef m;
    uint8_t *rows;  // SignLayout rows (item state)
    size_t row_pitch;
    size_t off_ints;  // byte offset of the int32 arrays (A, s1^, s2^, t0^) in a row
    size_t n;
    // Resident keys (optional): when key_idx is set, item i signs with key
    // k = min(key_idx[i], nkeys - 1); its secret key is sk + k * sk_pitch and
    // its expanded arrays (A, s1^, s2^, t0^) sit at key_ints + k * key_pitch
    // instead of in the item's row.  (Out-of-range indices are clamped so
    // that no access leaves the store.)
    const uint32_t *key_idx;
    const uint8_t *key_ints;
    size_t key_pitch;
    uint32_t nkeys;
    // Randomized signing (optional): item i's rhoprime is the 64 bytes at
    // rnd + i * rnd_pitch instead of SHAKE256(key || mu) (api.cu:143-148).
    const uint8_t *rnd;
    size_t rnd_pitch;
    // Long messages: mu / rhoprime with one sponge per 16-lane row
    // (sg_mu_row_kernel) instead of one lane per item.
    int row_mu;
};

__device__ __forceinline__ size_t sk_index(const SignArgs &a, size_t i) {
    if (!a.key_idx) return i;
    const uint32_t k = a.key_idx[i];
    return k < a.nkeys ? k : a.nkeys - 1;
}

// Scratch and signature destination of attempt slots.
struct SlotRef {
    uint8_t *base;  // SlotLayout of slot s at base + s * pitch
    size_t pitch;
    uint8_t *sig;  // signature bytes of slot s at sig + s * sig_pitch
    size_t sig_pitch;
};

template <class P>
struct ItemView {
    using Z = Sizes<P>;
    using SLy = SignLayout<P>;
    uint8_t *row;
    const int32_t *mat, *s1h, *s2h, *t0h;
    __device__ ItemView(const SignArgs &a, size_t item) {
        row = a.rows + item * a.row_pitch;
        mat = a.key_idx ? reinterpret_cast<const int32_t *>(a.key_ints + sk_index(a, item) * a.key_pitch)
                        : reinterpret_cast<const int32_t *>(row + a.off_ints);
        s1h = mat + P::K * P::L * N;
        s2h = s1h + P::L * N;
        t0h = s2h + P::K * N;
    }
};

__device__ __forceinline__ void load_words8(uint64_t (&w)[8], const uint8_t *p) {
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = reinterpret_cast<const uint64_t *>(p)[i];
}

// ---------------------------------------------------------------- fixed part

template <class P>
__device__ __forceinline__ void stage_mu(const SignArgs &a, size_t i) {
    using Z = Sizes<P>;
    using SLy = SignLayout<P>;
    const uint8_t *sk = a.sk + sk_index(a, i) * a.sk_pitch;
    uint64_t tr[4], key[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        tr[w] = load_le64(sk + Z::SK_TR + 8 * w);
        key[w] = load_le64(sk + Z::SK_KEY + 8 * w);
    }
    const uint8_t *m = a.m.ptr(i);
    const uint32_t mlen = a.m.len(i);
    Lane64 s[25];
    shake_absorb<SHAKE256_RATE>(s, [&](uint32_t w, uint32_t nb) -> uint64_t {
        if (w < 4) return w == 0 ? tr[0] : (w == 1 ? tr[1] : (w == 2 ? tr[2] : tr[3]));
        const uint8_t *p = m + 8 * (w - 4);
        return nb == 8 ? load_le64(p) : load_le_partial(p, nb);
    }, SEEDBYTES + mlen);
    uint64_t in[12];
    uint64_t mu[8];
    squeeze_words<8>(s, mu);
    uint64_t rp[8];
    if (a.rnd) {
        const uint8_t *r = a.rnd + i * a.rnd_pitch;
#pragma unroll
        for (int w = 0; w < 8; ++w) rp[w] = load_le64(r + 8 * w);
    } else {
#pragma unroll
        for (int w = 0; w < 4; ++w) in[w] = key[w];
#pragma unroll
        for (int w = 0; w < 8; ++w) in[4 + w] = mu[w];
        shake_absorb_regs<SHAKE256_RATE, SEEDBYTES + CRHBYTES>(s, in);
        squeeze_words<8>(s, rp);
    }
    uint64_t *dst = reinterpret_cast<uint64_t *>(a.rows + i * a.row_pitch + SLy::MU);
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        dst[w] = mu[w];
        dst[8 + w] = rp[w];
    }
}

// stage_mu with two lanes per item (keccak_f1600_pair): lane h of the pair
// holds the 32-bit halves h of the state.  A permutation takes 2/3 of the
// one-lane form's time for 1.2 x its instructions, so the serial absorb of
// the batch's longest message, which bounds the fixed stage, is shorter.
template <class P>
__device__ __forceinline__ void stage_mu_pair(const SignArgs &a, size_t i, uint32_t h) {
    using Z = Sizes<P>;
    using SLy = SignLayout<P>;
    const uint8_t *sk = a.sk + sk_index(a, i) * a.sk_pitch;
    const uint8_t *m = a.m.ptr(i);
    const uint32_t mlen = a.m.len(i);
    uint32_t s[25];
    shake_pair<SHAKE256_RATE>(s, h, SEEDBYTES + mlen, [&](uint32_t w) -> uint32_t {
        if (w < 4) return load_half_bytes(sk + Z::SK_TR, 8 * w + 4 * h, SEEDBYTES);
        return load_half_bytes(m, 8 * (w - 4) + 4 * h, mlen);
    });
    uint32_t *row = reinterpret_cast<uint32_t *>(a.rows + i * a.row_pitch);
#pragma unroll
    for (int w = 0; w < 8; ++w) row[SLy::MU / 4 + 2 * w + h] = s[w];
    if (a.rnd) {
        const uint8_t *r = a.rnd + i * a.rnd_pitch;
#pragma unroll
        for (int w = 0; w < 8; ++w) row[SLy::RHOPRIME / 4 + 2 * w + h] = load_half_bytes(r, 8 * w + 4 * h, CRHBYTES);
        return;
    }
    // rhoprime = SHAKE256(key || mu): 96 bytes, one block
    uint32_t t[25];
#pragma unroll
    for (int w = 0; w < 25; ++w) t[w] = 0u;
#pragma unroll
    for (int w = 0; w < 4; ++w) t[w] = load_half_bytes(sk + Z::SK_KEY, 8 * w + 4 * h, SEEDBYTES);
#pragma unroll
    for (int w = 0; w < 8; ++w) t[4 + w] = s[w];
    if (h == 0) t[12] ^= 0x1Fu;
    if (h == 1) t[SHAKE256_RATE / 8 - 1] ^= 0x80000000u;
    keccak_f1600_pair(t, h);
#pragma unroll
    for (int w = 0; w < 8; ++w) row[SLy::RHOPRIME / 4 + 2 * w + h] = t[w];
}

template <class P>
__device__ __forceinline__ void stage_expand_a(const SignArgs &a, size_t i, int k, int l) {
    const uint8_t *sk = a.sk + i * a.sk_pitch;
    uint64_t rho[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) rho[w] = load_le64(sk + 8 * w);
    int32_t *dst = reinterpret_cast<int32_t *>(a.rows + i * a.row_pitch + a.off_ints) + (k * P::L + l) * N;
    // Four consecutive coefficients per 16-byte store: lanes write different
    // rows, so each store instruction touches one cache line per lane, and a
    // store per coefficient made the kernel wait on its writes.  (Row pitches
    // and the int32 area offset are multiples of 16; the entry points check.)
    int4 *dst4 = reinterpret_cast<int4 *>(dst);
    int32_t b0 = 0, b1 = 0, b2 = 0;  // the last three accepted coefficients
    uniform_poly(rho, (uint32_t)((k << 8) | l), [&](int ctr, int32_t v) {
        if ((ctr & 3) == 3) dst4[ctr >> 2] = make_int4(b0, b1, b2, v);
        b0 = b1;
        b1 = b2;
        b2 = v;
    });
}

// poly p of the item: s1[0..L), s2[0..K), t0[0..K) -> NTT domain (layout D)
template <class P>
__device__ __forceinline__ void stage_prep(const SignArgs &a, size_t i, int p, int32_t *x, const ZetaLds &zt) {
    using Z = Sizes<P>;
    const uint8_t *sk = a.sk + i * a.sk_pitch;
    int32_t *dst = reinterpret_cast<int32_t *>(a.rows + i * a.row_pitch + a.off_ints) + (P::K * P::L + p) * N;
    int32_t v[4];
    uint32_t f[4];
    if (p < P::L + P::K) {
        wave_fields_A<Z::ETA_BITS>(f, sk + Z::SK_S1 + p * Z::POLYETA_BYTES);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = P::ETA - (int32_t)f[r];
    } else {
        wave_fields_A<13>(f, sk + Z::SK_T0 + (p - P::L - P::K) * POLYT0_BYTES);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = (1 << (D - 1)) - (int32_t)f[r];
    }
    ntt_wave(v, x, zt);
    wave_store_D(dst, v);
}

// ------------------------------------------------------------- attempt part

template <class P>
__device__ __forceinline__ void stage_y(const SignArgs &a, size_t item, uint32_t nonce, int l, uint8_t *slot) {
    using Z = Sizes<P>;
    uint64_t rp[8];
    load_words8(rp, a.rows + item * a.row_pitch + SignLayout<P>::RHOPRIME);
    mask_raw(rp, (uint32_t)(P::L * nonce + l) & 0xFFFF,
             reinterpret_cast<uint64_t *>(slot + SlotLayout<P>::YRAW + l * Z::YRAW_BYTES));
}

// stage_y with two lanes per sponge (both lanes of the pair call it)
template <class P>
__device__ __forceinline__ void stage_y_pair(const SignArgs &a, size_t item, uint32_t nonce, int l, uint8_t *slot,
                                             uint32_t h) {
    using Z = Sizes<P>;
    mask_raw_pair(reinterpret_cast<const uint32_t *>(a.rows + item * a.row_pitch + SignLayout<P>::RHOPRIME),
                  (uint32_t)(P::L * nonce + l) & 0xFFFF,
                  reinterpret_cast<uint32_t *>(slot + SlotLayout<P>::YRAW + l * Z::YRAW_BYTES), h);
}

template <class P>
__device__ __forceinline__ void load_y_A(int32_t (&v)[4], const uint8_t *slot, int l) {
    using Z = Sizes<P>;
    uint32_t f[4];
    wave_fields_A<Z::Z_BITS>(f, slot + SlotLayout<P>::YRAW + l * Z::YRAW_BYTES);
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = P::GAMMA1 - (int32_t)f[r];
}

// w for M attempts of the same item (slots[0..M)) in one wave: every A row
// is loaded once for all M products, and the M transforms give the wave
// independent work to overlap with the loads.
// ylds: with M > 1 and L = 7 the M x 7 transformed y polynomials live in
// this wave's LDS area (M * L * N words) instead of registers, where they
// would not fit beside the A rows and the INTT
template <class P, int M>
__device__ __forceinline__ void stage_w_multi(const SignArgs &a, size_t item, uint8_t *const (&slots)[M], int32_t *x,
                                              const ZetaLds &zt, int32_t *ylds = nullptr) {
    using Z = Sizes<P>;
    using SL = SlotLayout<P>;
// smallest L whose paired y vectors go to LDS (for L = 4 and 5 moving them
// there freed 10-27 VGPRs and changed nothing measurable)
#ifndef DIL_W_Y_LDS_MIN_L
#define DIL_W_Y_LDS_MIN_L 6
#endif
    constexpr bool Y_LDS = M > 1 && P::L >= DIL_W_Y_LDS_MIN_L;
    const ItemView<P> it(a, item);
    // A rows are streamed from HBM with the next row's loads issued before
    // the current row's products and INTTs (software prefetch)
    int32_t an[P::L][4];
#pragma unroll
    for (int l = 0; l < P::L; ++l) wave_load_D(an[l], it.mat + l * N);
    int32_t yh[Y_LDS ? 1 : M][Y_LDS ? 1 : P::L][4];
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int l = 0; l < P::L; ++l) {
            int32_t t[4];
            load_y_A<P>(t, slots[j], l);
            ntt_wave(t, x, zt);
            if constexpr (Y_LDS) {
                wave_store_D(ylds + (j * P::L + l) * N, t);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) yh[j][l][r] = t[r];
            }
        }
    for (int k = 0; k < P::K; ++k) {
        int32_t acc[M][4];
        int32_t ac[P::L][4];
#pragma unroll
        for (int l = 0; l < P::L; ++l)
#pragma unroll
            for (int r = 0; r < 4; ++r) ac[l][r] = an[l][r];
        if (k + 1 < P::K) {
#pragma unroll
            for (int l = 0; l < P::L; ++l) wave_load_D(an[l], it.mat + ((k + 1) * P::L + l) * N);
        }
#pragma unroll
        for (int j = 0; j < M; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[j][r] = 0;
#pragma unroll
            for (int l = 0; l < P::L; ++l) {
                if constexpr (Y_LDS) {
                    int32_t yv[4];
                    wave_load_D(yv, ylds + (j * P::L + l) * N);
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[j][r] += mont_mul(ac[l][r], yv[r]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[j][r] += mont_mul(ac[l][r], yh[j][l][r]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < M; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[j][r] = reduce32(acc[j][r]);
            intt_wave(acc[j], x, zt);
            int32_t w1[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[j][r] = caddq(acc[j][r]);
                int32_t w0;
                w1[r] = decompose<P>(acc[j][r], w0);
            }
            wave_store_A(reinterpret_cast<int32_t *>(slots[j] + SL::W) + k * N, acc[j]);
            layout_A_to_D(w1, x);
            uint32_t u[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) u[r] = (uint32_t)w1[r];
            wave_pack_D<Z::W1_BITS>(slots[j] + SL::W1PK + k * Z::POLYW1_BYTES, u);
        }
    }
}

template <class P>
__device__ __forceinline__ void stage_w(const SignArgs &a, size_t item, uint8_t *slot, int32_t *x, const ZetaLds &zt) {
    uint8_t *const slots[1] = {slot};
    stage_w_multi<P, 1>(a, item, slots, x, zt);
}

template <class P>
__device__ __forceinline__ void stage_c(const SignArgs &a, size_t item, uint8_t *slot, uint8_t *sig_out) {
    using Z = Sizes<P>;
    using SL = SlotLayout<P>;
    const uint64_t *mu = reinterpret_cast<const uint64_t *>(a.rows + item * a.row_pitch + SignLayout<P>::MU);
    const uint64_t *w1 = reinterpret_cast<const uint64_t *>(slot + SL::W1PK);
    Lane64 s[25];
    shake_absorb<SHAKE256_RATE>(s, [&](uint32_t w, uint32_t) { return w < 8 ? mu[w] : w1[w - 8]; },
                                (uint32_t)Z::CHASH_IN_BYTES);
    uint64_t ct[4];
    squeeze_words<4>(s, ct);
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
        for (int b = 0; b < 8; ++b) sig_out[8 * w + b] = (uint8_t)(ct[w] >> (8 * b));
    challenge<P::TAU>(ct, reinterpret_cast<uint32_t *>(slot + SL::CP));
}

// stage_c over a lane pair (keccak_f1600_pair): lanes 2j and 2j + 1 hash the
// same slot, each holding one 32-bit half of every state word, so a wave
// runs the challenge hash of 32 slots with a chain a third shorter than one
// state per lane.  h = lane & 1; both lanes of the pair must call it.
template <class P>
__device__ __forceinline__ void stage_c_pair(const SignArgs &a, size_t item, uint8_t *slot, uint8_t *sig_out,
                                             uint32_t h) {
    using Z = Sizes<P>;
    using SL = SlotLayout<P>;
    const uint32_t *mu = reinterpret_cast<const uint32_t *>(a.rows + item * a.row_pitch + SignLayout<P>::MU);
    const uint32_t *w1 = reinterpret_cast<const uint32_t *>(slot + SL::W1PK);
    uint32_t s[25];
    shake_absorb_pair<SHAKE256_RATE, Z::CHASH_IN_BYTES>(
        s, [&](uint32_t w) { return w < 8 ? mu[2 * w + h] : w1[2 * (w - 8) + h]; }, h);
    keccak_f1600_pair(s, h);
    uint64_t ct[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) ct[w] = pair_word(s, w, h);
    // c~: each lane of the pair writes its 4-byte halves (16 bytes)
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t v = s[w];
#pragma unroll
        for (int b = 0; b < 4; ++b) sig_out[8 * w + 4 * h + b] = (uint8_t)(v >> (8 * b));
    }
    // the challenge block squeezed in the pair layout as well (one state per
    // lane: ~1 % slower loop for Dilithium3)
#ifndef DIL_CHALLENGE_PAIR
#define DIL_CHALLENGE_PAIR 1
#endif
    if constexpr (DIL_CHALLENGE_PAIR)
        challenge_pair<P::TAU>(ct, reinterpret_cast<uint32_t *>(slot + SL::CP), h);
    else
        challenge<P::TAU>(ct, reinterpret_cast<uint32_t *>(slot + SL::CP), h == 0);
}

// Returns true (wave-uniform) when the attempt passes every check; then
// sig_out holds the complete signature.  An attempt is accepted iff all of
// ||r0|| < gamma2 - beta, ||c t0|| < gamma2, #hints <= omega and
// ||z|| < gamma1 - beta hold (sign.cu:986-1196), so the order of the checks
// does not change any result; the r0 check runs first because it rejects
// most often (about 2/3 of Dilithium3 attempts), then c t0 and the hints,
// then z.  Fields of a rejected attempt's sig_out are left partly written.
template <class P>
__device__ __forceinline__ bool stage_z(const SignArgs &a, size_t item, uint8_t *slot, uint8_t *sig_out, int32_t *x,
                                        const ZetaLds &zt) {
    using Z = Sizes<P>;
    using SL = SlotLayout<P>;
    const ItemView<P> it(a, item);
    const int lane = lane_id();
    int32_t ch[4];
    wave_load_challenge_A(ch, reinterpret_cast<const uint32_t *>(slot + SL::CP));
    ntt_wave(ch, x, zt);

    uint8_t *h = sig_out + Z::SIG_H;
    int pos = 0;
    // operands of row k + 1 are loaded while row k is transformed (prefetch)
    const int32_t *wbase = reinterpret_cast<const int32_t *>(slot + SL::W);
    int32_t nr0[4], nct0[4], nw[4];
    wave_load_D(nr0, it.s2h);
    wave_load_D(nct0, it.t0h);
    wave_load_A(nw, wbase);
    for (int k = 0; k < P::K; ++k) {
        int32_t r0[4], ct0[4], w[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            r0[r] = nr0[r];
            ct0[r] = nct0[r];
            w[r] = nw[r];
        }
        if (k + 1 < P::K) {
            wave_load_D(nr0, it.s2h + (k + 1) * N);
            wave_load_D(nct0, it.t0h + (k + 1) * N);
            wave_load_A(nw, wbase + (k + 1) * N);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            r0[r] = mont_mul(ch[r], r0[r]);
            ct0[r] = mont_mul(ch[r], ct0[r]);
        }
        intt_wave(r0, x, zt);
        int32_t w1[4];
        bool big = false;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int32_t w0;
            w1[r] = decompose<P>(w[r], w0);
            r0[r] = reduce32(w0 - r0[r]);
            big |= exceeds(r0[r], P::GAMMA2 - P::BETA);
        }
        if (wave_any(big)) return false;
        intt_wave(ct0, x, zt);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            ct0[r] = reduce32(ct0[r]);
            big |= exceeds(ct0[r], P::GAMMA2);
        }
        if (wave_any(big)) return false;
        // MakeHint in ascending coefficient order j = 64 r + lane
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int hb = make_hint<P>(r0[r] + ct0[r], w1[r]);
            const uint64_t m = __ballot(hb);
            const int cnt = __popcll(m);
            if (pos + cnt > P::OMEGA) return false;
            if (hb) {
                const int idx = pos + __popcll(m & ((1ull << lane) - 1ull));
                h[idx] = (uint8_t)(64 * r + lane);
            }
            pos += cnt;
        }
        if (lane == 0) h[P::OMEGA + k] = (uint8_t)pos;
    }

    for (int l = 0; l < P::L; ++l) {
        int32_t zl[4], s[4], y[4];
        wave_load_D(s, it.s1h + l * N);
        load_y_A<P>(y, slot, l);
#pragma unroll
        for (int r = 0; r < 4; ++r) zl[r] = mont_mul(ch[r], s[r]);
        intt_wave(zl, x, zt);
        bool big = false;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            zl[r] = reduce32(zl[r] + y[r]);
            big |= exceeds(zl[r], P::GAMMA1 - P::BETA);
        }
        if (wave_any(big)) return false;
        layout_A_to_D(zl, x);
        uint32_t f[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) f[r] = (uint32_t)(P::GAMMA1 - zl[r]);
        wave_pack_D<Z::Z_BITS>(sig_out + Z::SIG_Z + l * Z::POLYZ_BYTES, f);
    }
    for (int j = pos + lane; j < P::OMEGA; j += 64) h[j] = 0;
    return true;
}

}  // namespace dil
