// Byte layout of one item's row in the device pools.
//
// The drop-in API receives caller-allocated pitched pools whose row sizes the
// caller computes with the reference's formulas (tests/test_vectors.cu:64-70,
// 81-100, 117-128, 149-166).  Every layout below fits inside those rows for
// every mode and message length (checked by layout_fits_reference_rows in the
// C-ABI and by tests/test_capi_symbols.py), so an existing caller's pools work
// unchanged.  The device-resident API allocates rows of exactly these sizes.
#pragma once

#include "params.hpp"

namespace dil {

template <class P>
struct KeypairLayout {
    using Z = Sizes<P>;
    static constexpr size_t PK = 0;
    static constexpr size_t SK = align256(Z::PK_BYTES);
    static constexpr size_t SCRATCH = SK + align256(Z::SK_BYTES);
    // scratch: rhoprime (64 B) then s1 | s2 as int8 polynomials
    static constexpr size_t RHOPRIME = SCRATCH;
    static constexpr size_t S8 = SCRATCH + 64;
    static constexpr size_t ROW = SCRATCH + (size_t)P::K * N * 4;
    static_assert(64 + (P::L + P::K) * N <= P::K * N * 4, "keypair scratch fits the reference row");
    // reference formula (test_vectors.cu:64-70)
    static constexpr size_t REF_ROW = align256(Z::PK_BYTES) + align256(Z::SK_BYTES) + (size_t)P::K * N * 4;
};

// Scratch of one signing attempt ("slot"): written by the y / w / c / z
// stages of one (item, nonce) attempt.
template <class P>
struct SlotLayout {
    using Z = Sizes<P>;
    static constexpr size_t W1PK = 0;                                 // K * POLYW1_BYTES
    static constexpr size_t CP = align256(Z::W1PK_BYTES);             // challenge masks (64 B)
    static constexpr size_t FLAG = CP + 256;                          // int32 flags
    static constexpr size_t YRAW = FLAG + 256;                        // L * 680 bytes
    static constexpr size_t W = YRAW + align256((size_t)P::L * Z::YRAW_BYTES);  // int32 K x 256
    static constexpr size_t BYTES = W + (size_t)P::K * N * 4;
};

template <class P>
struct SignLayout {
    using Z = Sizes<P>;
    using SL = SlotLayout<P>;
    static constexpr size_t SIG = 0;
    static constexpr size_t SK = align256(Z::SIG_BYTES);
    static constexpr size_t MU = SK + align256(Z::SK_BYTES);  // mu (64) | rhoprime (64) | attempts (4)
    static constexpr size_t RHOPRIME = MU + CRHBYTES;
    static constexpr size_t ATTEMPTS = MU + 2 * CRHBYTES;
    static constexpr size_t SLOT = MU + 256;
    static constexpr size_t M = SLOT + SL::BYTES;
    static size_t ints(size_t mlen) { return M + align256(mlen); }
    // int32 arrays after the message: A (K*L), s1hat (L), s2hat (K), t0hat (K)
    static size_t mat(size_t mlen) { return ints(mlen); }
    static size_t s1hat(size_t mlen) { return mat(mlen) + (size_t)P::K * P::L * N * 4; }
    static size_t s2hat(size_t mlen) { return s1hat(mlen) + (size_t)P::L * N * 4; }
    static size_t t0hat(size_t mlen) { return s2hat(mlen) + (size_t)P::K * N * 4; }
    static size_t row(size_t mlen) { return t0hat(mlen) + (size_t)P::K * N * 4; }
    // reference formula (test_vectors.cu:81-100)
    static size_t ref_row(size_t mlen) {
        return align256(Z::SIG_BYTES) + align256(Z::SK_BYTES) + align256(SEEDBYTES) + align256(SEEDBYTES + mlen) +
               align256(SEEDBYTES + CRHBYTES + Z::W1PK_BYTES) + align256(CRHBYTES) +
               (size_t)(P::K * P::L + 3 * P::L + 4 * P::K + 1) * N * 4;
    }
};

template <class P>
struct TempLayout {
    using Z = Sizes<P>;
    static constexpr size_t SIG = 0;
    static constexpr size_t SLOT = align256(Z::SIG_BYTES);
    static constexpr size_t ROW = SLOT + SlotLayout<P>::BYTES;
    // reference formula (test_vectors.cu:117-128)
    static constexpr size_t REF_ROW = align256(Z::SIG_BYTES) + align256(CRHBYTES + Z::W1PK_BYTES) +
                                      (size_t)(2 * P::L + 3 * P::K) * N * 4;
    static_assert(ROW <= REF_ROW, "temp row fits the reference row");
};

template <class P>
struct VerifyLayout {
    using Z = Sizes<P>;
    static constexpr size_t SIG = 0;
    static constexpr size_t PK = align256(Z::SIG_BYTES);
    static constexpr size_t HASH = PK + align256(Z::PK_BYTES);  // mu | w1' packed
    static constexpr size_t CP = HASH + align256(Z::CHASH_IN_BYTES);  // challenge masks (64 B)
    static constexpr size_t HINT = CP + 256;                          // K x 256 bits
    static constexpr size_t FLAG = HINT + 256;                        // int32 status
    static constexpr size_t M = FLAG + 256;
    static size_t row(size_t mlen) { return M + align256(mlen); }
    // The verify kernels touch only [HASH, M) of a work row when signature,
    // public key and message live elsewhere (the device-resident entry
    // points): a compact work area of WORK_BYTES per item serves as rows based
    // HASH bytes before it.
    static constexpr size_t WORK_BEGIN = HASH;
    static constexpr size_t WORK_BYTES = M - HASH;
    // reference formula (test_vectors.cu:149-166)
    static size_t ref_row(size_t mlen) {
        return align256(Z::SIG_BYTES) + align256(Z::PK_BYTES) + align256(SEEDBYTES + mlen) +
               align256(CRHBYTES + Z::W1PK_BYTES) + align256(SEEDBYTES) +
               (size_t)(P::K * P::L + P::L + 3 * P::K + 1) * N * 4 + 4;
    }
};

// FLAGS word of a resident key: set when the key was loaded or generated
// with its secret key (verify-only keys hold zeros in SK, s1^, s2^, t0^)
constexpr uint32_t KEY_HAS_SECRET = 1u;

// One resident key of a key store: public key, secret key (absent for
// verify-only keys), tr = H(pk), and the expanded arrays every signature and
// verification under the key would otherwise recompute: A = ExpandA(rho)
// (K*L polys, 5*K*L Keccak permutations), s1^, s2^, t0^ (NTT domain) and
// t1^ = NTT(t1 * 2^d), int32 in the NTT kernels' register layout.
template <class P>
struct KeyLayout {
    using Z = Sizes<P>;
    static constexpr size_t PK = 0;
    static constexpr size_t SK = align256(Z::PK_BYTES);
    static constexpr size_t TR = SK + align256(Z::SK_BYTES);  // tr (32) | flags (4)
    static constexpr size_t FLAGS = TR + SEEDBYTES;           // KEY_HAS_SECRET or 0 (verify-only)
    static constexpr size_t INTS = TR + 256;
    static constexpr size_t MAT = INTS;
    static constexpr size_t S1H = MAT + (size_t)P::K * P::L * N * 4;
    static constexpr size_t S2H = S1H + (size_t)P::L * N * 4;
    static constexpr size_t T0H = S2H + (size_t)P::K * N * 4;
    static constexpr size_t T1H = T0H + (size_t)P::K * N * 4;
    static constexpr size_t ROW = T1H + (size_t)P::K * N * 4;
    // keygen into the store uses the (not yet expanded) int area as scratch
    static_assert(KeypairLayout<P>::ROW - KeypairLayout<P>::SCRATCH <= ROW - INTS, "keygen scratch fits");
};

}  // namespace dil
