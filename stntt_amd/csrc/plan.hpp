// plan.hpp -- how a transform of size 2^L is split into LDS-resident passes.
//
// The reference runs one global-memory launch per radix-2 stage until the
// remaining stages fit one block (src/cuda/ntt/implementations/
// st_ntt_global_radix2.cuh:98-116): 14 HBM passes at n = 2^24.  Here a
// small transform is one LDS pass, and a large one is 2-3 tile passes of up
// to 9 stages each with whole 128-byte rows moved per access: n = 2^24 and
// n = 2^26 take 3 passes.
#pragma once

#include <type_traits>

#include "ntt_big.hpp"
#include "ntt_pass.hpp"
#include "ntt_tile.hpp"

namespace stntt {

constexpr int PASS_THREADS = 256;
constexpr int MAX_LOGR = 14;        // largest pass radix; also the size of the hi twiddle factor
constexpr int MIN_LOGR_TILE = 4;    // smallest radix of a tile pass
constexpr int MAX_LOGR_TILE = 9;    // largest radix of a tile pass (64 KiB tile)
constexpr int MIN_LOGR_BIG = 10;    // passes of radix 2^10..2^12 run the large-radix kernel
constexpr int MAX_LOGR_BIG = 12;
// Single-pass sizes with a one-transform variant (a workgroup of
// narrow_threads per transform): up to 2^10 for 8-byte words, 2^12 for
// 4-byte words.  A lone transform in one pass against two tile passes:
// Goldilocks 2^9 7.0 against 10.3 us, 2^10 9.7 against 10.7 us (2^11 12.2
// against 11.4, 2^12 13.9 against 11.4: two passes win); 32-bit 2^10 8.1
// against 9.3 us, 2^11 8.8 against 10.7, 2^12 9.5 against 9.8 us.
template <class FP>
constexpr int narrow_max_logr() {
    return sizeof(typename FP::T) == 4 ? 12 : 10;
}
// threads of a one-transform single pass: 8 words per thread for 8-byte
// words, 16 for 4-byte words, at least one wave
template <class FP, int LOGR>
constexpr int narrow_threads() {
    constexpr int ept = sizeof(typename FP::T) == 4 ? 16 : 8;
    return (1 << LOGR) / ept < 64 ? 64 : (1 << LOGR) / ept;
}

// log2 of the tile size in elements: 64 KiB of LDS per workgroup
template <class FP>
constexpr int tile_log() {
    return sizeof(typename FP::T) == 8 ? 13 : 14;  // 64 KiB of words
}

template <class FP>
constexpr int MAX_LOGR_SINGLE = tile_log<FP>();

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// columns per tile: whole transforms for single-pass tiles (up to 256),
// consecutive t = sJ + j for multi-pass tiles (up to 16)
template <class FP>
constexpr int pass_logc(int logr, bool single) {
    return single ? cmax(0, cmin(tile_log<FP>() - logr, 8)) : cmax(0, cmin(tile_log<FP>() - logr, 4));
}

struct Schedule {
    bool single = true;
    int npasses = 0;
    int logr[4] = {0, 0, 0, 0};
    int logc[4] = {-1, -1, -1, -1};  // large-radix passes: tile columns (log2), -1 = big_logc_choice
};

// Columns (log2) of a large-radix pass over `batches` transforms of 2^L
// words, or -1 when no tile shape fits: at most the full 128 KiB tile and at
// least a quarter of it, with n / R >= C columns and, on a later pass,
// J >= C so the columns of a tile share their pass twiddles.  `want` >= 0
// asks for that width; otherwise the widest tile that still gives every CU
// two tiles, else the narrowest that fits.
template <class FP>
inline int big_logc_choice(int L, int logr, int log_j, bool first, uint64_t batches, int num_cus, int want = -1) {
    const int full = big_full_logc<FP>(logr);
    auto fits = [&](int lc) { return lc >= full - 2 && lc <= full && L - logr >= lc && (first || log_j >= lc); };
    if (want >= 0) return fits(want) ? want : -1;
    int last = -1;
    for (int lc = full; lc >= full - 2; --lc) {
        if (!fits(lc)) continue;
        last = lc;
        if ((batches << (L - logr)) >> lc >= 2ull * (uint64_t)num_cus) return lc;
    }
    return last;
}

// smallest L a tile schedule covers: every pass needs n / R >= C columns and
// every later pass J >= C
template <class FP>
constexpr int min_tile_log() {
    return 2 * tile_logc<FP>();
}

// Single LDS pass when the transform is small or the batch alone fills the
// GPU; otherwise 2-3 tile passes of at most 2^9 points each (2^24: 8+8+8).
// Sizes whose default schedule is two large-radix passes (2^10..2^12 each).
constexpr bool BIG_BY_DEFAULT = false;

template <class FP>
inline Schedule make_schedule(int L, uint64_t batches, bool allow_big = true) {
    Schedule s;
    const bool fits_one = L <= MAX_LOGR_SINGLE<FP>;
    // Sizes up to narrow_max_logr always run in one pass: below the two-pass
    // tile minimum there is no other schedule, and above it one pass beat
    // two tile passes (measured above).
    if (L < min_tile_log<FP>() || L <= narrow_max_logr<FP>() ||
        (fits_one && (batches << L) >= (1ull << 20) && batches >= 128)) {
        s.single = true;
        s.npasses = 1;
        s.logr[0] = L;
        return s;
    }
    s.single = false;
    if (BIG_BY_DEFAULT && allow_big && L >= 2 * MIN_LOGR_BIG && L <= 2 * MAX_LOGR_BIG) {
        s.npasses = 2;
        s.logr[0] = L / 2;
        s.logr[1] = L - L / 2;
        return s;
    }
    int np = (L + MAX_LOGR_TILE - 1) / MAX_LOGR_TILE;
    if (np < 2) np = 2;
    s.npasses = np;
    for (int i = 0; i < np; ++i) s.logr[i] = L / np + (i < L % np ? 1 : 0);
    return s;
}

}  // namespace stntt
