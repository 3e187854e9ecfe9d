// plan.hpp -- how a transform of size 2^L is split into LDS-resident passes.
//
// The reference runs one global-memory launch per radix-2 stage until the
// remaining stages fit one block (src/cuda/ntt/implementations/
// st_ntt_global_radix2.cuh:98-116): 14 HBM passes at n = 2^24.  Here each
// pass consumes up to 12-14 stages from a tile held in LDS, so n = 2^24 takes
// 2 passes and n = 2^26 takes 3.
#pragma once

#include <type_traits>

#include "ntt_pass.hpp"

namespace stntt {

constexpr int PASS_THREADS = 256;
constexpr int MAX_LOGR = 14;        // largest pass radix; also the size of the hi twiddle factor
constexpr int MIN_LOGR_MULTI = 6;   // smallest radix a multi-pass schedule uses
constexpr int MAX_LOGR_MULTI = 12;  // largest radix a multi-pass schedule uses

// log2 of the tile size in elements: 64 KiB of LDS per workgroup
template <class FP>
constexpr int tile_log() {
    return std::is_same_v<FP, FieldGL> ? 13 : 14;
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
};

template <class FP>
inline Schedule make_schedule(int L) {
    Schedule s;
    if (L <= MAX_LOGR_SINGLE<FP>) {
        s.single = true;
        s.npasses = 1;
        s.logr[0] = L;
        return s;
    }
    s.single = false;
    int np = (L + MAX_LOGR_MULTI - 1) / MAX_LOGR_MULTI;
    if (np < 2) np = 2;
    s.npasses = np;
    for (int i = 0; i < np; ++i) s.logr[i] = L / np + (i < L % np ? 1 : 0);
    return s;
}

}  // namespace stntt
