// Keccak-f[1600] with one state per 16-lane DPP row, for long serial sponges.
//
// A message of L bytes costs floor(L / 136) + 1 dependent permutations in a
// sponge, whatever the batch: a 64 KiB message is 482 of them.  With one
// state per lane (keccak.hpp) each permutation is ~4,300 dependent VALU
// instructions of one lane; here ten lanes share a state and a round is 54
// VALU instructions and one LDS round trip.  A lone wavefront runs a chained
// permutation in 3.45 us on MI355X, against 8.64 us with one state per lane
// and 5.93 us with two lanes per state (keccak_f1600_pair); the price is
// throughput, 2.1 G permutations/s for the whole card against 8.8 G:
//
//   lane g = 2 y + h of the row (g < 10) holds half h (h = 0: bits 0-31,
//   h = 1: bits 32-63) of the five words (x, y), x = 0..4, of row y in five
//   VGPRs; lanes 10..15 of the row hold zeros
//
//   theta  column parities: XOR over the row's lanes of equal h with three
//          DPP row rotations (by 2, 4, 8; the zero lanes pad the 16-lane row);
//          rotl(C, 1) takes the other half from the partner lane (quad_perm)
//   rho    per lane and word one funnel shift of the own half (high operand)
//          with the partner's (low operand): for R > 32 and R = 0 the lane
//          produces the pair's other half of the rotated word instead, which
//          pi writes to that half's place, so no lane selects operands
//   pi     (x, y) -> (y, 2x + 3y) moves words across rows: five LDS writes to
//          per-lane addresses and one row read back (in-order LDS traffic of
//          one wavefront); padding lanes write to a dummy word and read zeros
//   chi    along x, inside the lane; iota on the y = 0 lanes
//
// A wavefront holds four states (DPP rows 0..3).  All 64 lanes must execute
// every call (DPP reads partner lanes): groups without work run on zeros.
// The round constants and rotation offsets are those of keccak.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "keccak.hpp"
#include "ntt.hpp"  // wave_lds_sync, lane_id

namespace dil {

// LDS words per state: the 50 halves of the pi exchange (index h * 25 +
// row * 5 + x), 5 zero words that padding lanes read, 1 dummy word that they
// write; the stride spreads the four states of a wave over different banks
constexpr int ROW_LDS_WORDS = 72;

// v ^ (v moved by a DPP row pattern): update_dpp with a zero old value
// folds into one v_xor_b32_dpp
template <int CTRL>
__device__ __forceinline__ uint32_t xor_dpp(uint32_t v) {
    return v ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
constexpr int DPP_ROW_ROR2 = 0x122, DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;

struct RowLane {
    uint32_t h, y;        // half, row (y >= 5: padding lane)
    bool active;          // y < 5
    uint32_t sh[5];       // rho funnel-shift amount of word (x, y)
    uint32_t waddr[5];    // pi destination of word x's rotated half (index into the LDS array)
    uint32_t raddr;       // this lane's row after pi (5 words)
    uint32_t iota_mask;   // all ones on the y = 0 lanes
};

// lds: this workgroup's LDS array; the wavefront uses 4 * ROW_LDS_WORDS words
// from wave_base.  Padding reads must see zeros, which this call writes
// (ordered before their first read by the first round's wave_lds_sync).
__device__ __forceinline__ RowLane row_lane_setup(uint32_t *lds, uint32_t wave_base) {
    RowLane r;
    const uint32_t lane = (uint32_t)lane_id(), g = lane & 15;
    const uint32_t st = wave_base + (lane >> 4) * ROW_LDS_WORDS;
    r.h = g & 1;
    r.y = g >> 1;
    r.active = r.y < 5;
    if (g < 5) lds[st + 50 + g] = 0u;
#pragma unroll
    for (int x = 0; x < 5; ++x) {
        const int R = r.active ? kKeccak.rho[x + 5 * (int)r.y] : 0;
        // half h of rotl(w, R) = funnel(own, other, 32 - R) for 0 < R <= 32;
        // for R > 32 and R = 0, rotl(w, R) = rotl(swap(w), R - 32) and the
        // same funnel with shift (64 - R) mod 32 gives half 1 - h of it
        const bool swapped = R == 0 || R > 32;
        r.sh[x] = (uint32_t)(swapped ? (64 - R) & 31 : 32 - R);
        r.waddr[x] = r.active ? st + (r.h ^ (uint32_t)swapped) * 25 + ((2 * x + 3 * r.y) % 5) * 5 + r.y : st + 55;
    }
    r.raddr = r.active ? st + r.h * 25 + r.y * 5 : st + 50;
    r.iota_mask = (r.active && r.y == 0) ? 0xFFFFFFFFu : 0u;
    return r;
}

__device__ __forceinline__ void keccak_round_row(uint32_t (&a)[5], const RowLane &r, uint32_t *lds,
                                                 uint32_t rc_half) {
    // column parities: XOR over the row's lanes of equal h, the five columns
    // interleaved so that no DPP read waits on the previous write
    uint32_t c[5], co[5];
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = xor_dpp<DPP_ROW_ROR2>(a[x]);
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = xor_dpp<DPP_ROW_ROR4>(c[x]);
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = xor_dpp<DPP_ROW_ROR8>(c[x]);
#pragma unroll
    for (int x = 0; x < 5; ++x) co[x] = pair_other(c[x]);
#pragma unroll
    for (int x = 0; x < 5; ++x)
        a[x] = bitop3<kXor3>(a[x], c[(x + 4) % 5], rotl_half<1>(c[(x + 1) % 5], co[(x + 1) % 5]));
    uint32_t o[5];
#pragma unroll
    for (int x = 0; x < 5; ++x) o[x] = pair_other(a[x]);
#pragma unroll
    for (int x = 0; x < 5; ++x) lds[r.waddr[x]] = funnel(a[x], o[x], r.sh[x]);
    wave_lds_sync();
    uint32_t b[5];
#pragma unroll
    for (int x = 0; x < 5; ++x) b[x] = lds[r.raddr + x];
    wave_lds_sync();  // reads done before the next round's writes
#pragma unroll
    for (int x = 0; x < 5; ++x) a[x] = bitop3<kChi>(b[x], b[(x + 1) % 5], b[(x + 2) % 5]);
    a[0] ^= rc_half & r.iota_mask;
}

// Fully unrolled: the round constants become immediates (a scalar load of
// the table per round would be waited for together with the LDS traffic,
// which shares its counter), and the short code of the row form keeps the
// kernel small.
template <int I = 0>
__device__ __forceinline__ void keccak_rounds_row(uint32_t (&a)[5], const RowLane &r, uint32_t *lds) {
    constexpr uint64_t rc = kKeccak.rc[I];
    keccak_round_row(a, r, lds, r.h ? (uint32_t)(rc >> 32) : (uint32_t)rc);
    if constexpr (I + 1 < 24) keccak_rounds_row<I + 1>(a, r, lds);
}

__device__ __forceinline__ void keccak_f1600_row(uint32_t (&a)[5], const RowLane &r, uint32_t *lds) {
    keccak_rounds_row(a, r, lds);
}

// Word index of (x, y) in the state; words >= RW are outside the rate.
__device__ __forceinline__ uint32_t row_word(const RowLane &r, int x) { return (uint32_t)x + 5u * r.y; }

// Absorbs a byte stream of `len` bytes (SHAKE256 domain byte 0x1F) and
// applies the first squeeze permutation: get_half(i) returns half r.h of
// input word i (bytes past the end zero; only called for i < ceil(len / 8)).
// Afterwards a[x] holds half h of output word x + 5 y of the first block.
template <int RATE, class GetHalf>
__device__ __forceinline__ void shake_row(uint32_t (&a)[5], const RowLane &r, uint32_t *lds, uint32_t len,
                                          GetHalf get_half) {
    constexpr uint32_t RW = RATE / 8;
#pragma unroll
    for (int x = 0; x < 5; ++x) a[x] = 0u;
    const uint32_t nwords = (len + 7) / 8, nfull = len / RATE;
    auto fetch = [&](uint32_t blk, int x) -> uint32_t {
        const uint32_t w = row_word(r, x);
        const uint32_t i = blk * RW + w;
        return (r.active && w < RW && i < nwords) ? get_half(i) : 0u;
    };
    uint32_t cur[5];
#pragma unroll
    for (int x = 0; x < 5; ++x) cur[x] = fetch(0, x);
#pragma unroll 1
    for (uint32_t blk = 0; blk < nfull; ++blk) {
#pragma unroll
        for (int x = 0; x < 5; ++x) a[x] ^= cur[x];
#pragma unroll
        for (int x = 0; x < 5; ++x) cur[x] = fetch(blk + 1, x);  // overlaps the permutation
        keccak_f1600_row(a, r, lds);
    }
    // last (partial) block and padding: 0x1F at byte rem, 0x80 at byte RATE - 1
    const uint32_t rem = len - nfull * RATE;
#pragma unroll
    for (int x = 0; x < 5; ++x) {
        const uint32_t w = row_word(r, x);
        uint32_t v = cur[x];
        if (r.active && w == rem / 8 && r.h == (rem % 8) / 4) v ^= 0x1Fu << (8 * (rem % 4));
        if (r.active && w == RW - 1 && r.h == 1) v ^= 0x80000000u;
        a[x] ^= v;
    }
    keccak_f1600_row(a, r, lds);
}

}  // namespace dil
