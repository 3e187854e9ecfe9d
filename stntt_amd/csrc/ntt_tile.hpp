// ntt_tile.hpp -- the large-n pass: a radix-R Stockham step over 128-byte
// column tiles, two register sub-steps and one LDS exchange.
//
// A pass applies (see ntt_pass.hpp for the stage algebra, restating
// src/ntt/ntt_cpu.cpp:66-79)
//     y[(R s + k) J + j] = w^(s J k) * sum_m x[s J + j + m n/R] w_R^(m k)
// to a tile of C consecutive t = s J + j ("columns") and all R rows.  C is
// chosen so a row segment is 128 bytes (16 Goldilocks words, 32 32-bit
// words): every global load and store of the pass moves whole cache lines,
// where the reference moves one 4-byte word per lane per stage
// (src/cuda/ntt/implementations/st_ntt_global_radix2.cuh:41-55).
//
// R = r1 * r2 with r1 = 2^ceil(LOGR/2):
//   sub-step 1: each thread owns one column c and one group g < R/r1; it
//     loads rows g + m R/r1 (m < r1) straight from HBM into registers, runs
//     an r1-point DFT, multiplies output k by w_R^(g k) and writes row
//     r1 g + k of the LDS tile;
//   sub-step 2: groups g2 < r1 read rows g2 + m r1 (m < r2) from LDS, run an
//     r2-point DFT and store output k = k2 r1 + g2, times the pass twiddle,
//     straight to HBM.
// The LDS tile is XOR-swizzled (column c of row q lives at c ^ (q mod C)),
// so both the column-major write and either read order are conflict-free.
#pragma once

#include "ntt_pass.hpp"

namespace stntt {

template <class FP>
constexpr int tile_logc() {
    return std::is_same_v<FP, FieldGL> ? 4 : 5;  // 128-byte rows
}

template <int LOGR>
constexpr int tile_lr1() {
    return (LOGR + 1) / 2;
}

template <class FP, int LOGR>
constexpr int tile_threads() {
    return (1 << (LOGR - tile_lr1<LOGR>())) << tile_logc<FP>();  // r2 * C
}

// J1: first pass of a schedule (J = 1): outputs of a column are contiguous
// (y[R t + k]); otherwise J >= C and a row of C outputs is contiguous.
template <class FP, int LOGR, int DIR, bool J1>
__global__ void __launch_bounds__((tile_threads<FP, LOGR>())) ntt_tile_kernel(PassArgs<FP> a) {
    using T = typename FP::T;
    constexpr int LOGC = tile_logc<FP>();
    constexpr int C = 1 << LOGC;
    constexpr int R = 1 << LOGR;
    constexpr int LR1 = tile_lr1<LOGR>(), LR2 = LOGR - LR1;
    constexpr int r1 = 1 << LR1, r2 = 1 << LR2;
    constexpr int NT = tile_threads<FP, LOGR>();
    constexpr int G2 = r1 / r2;  // sub-step-2 groups per thread (1 or 2)
    static_assert(NT == (R / r1) * C, "one sub-step-1 group per thread");

    __shared__ T lds[R * C];
    __shared__ typename FP::Tw twl[J1 ? 1 : R];  // shared pass twiddles w^(s J k), k < R

    const uint32_t tid = threadIdx.x;
    const uint32_t logn = a.log_n;
    const uint32_t tpt = a.tiles_per_transform;
    const uint32_t b = blockIdx.x / tpt;
    const uint32_t t0 = (blockIdx.x % tpt) << LOGC;  // first column t of the tile
    const uint64_t base = (uint64_t)b << logn;
    const uint32_t stride_in = 1u << (logn - LOGR);  // n / R
    const T *in = a.in + base;
    T *out = a.out + base;

    // ---- pass twiddles shared by the tile (all columns have the same s) ----
    const uint32_t sJ = J1 ? 0u : ((t0 >> a.log_j) << a.log_j);
    if constexpr (!J1) {
        if (sJ != 0) {
            for (uint32_t k = tid; k < (uint32_t)R; k += NT) {
                const uint32_t e = sJ * k;  // < n
                const uint32_t log_b = logn - a.log_rmax;
                twl[k] = FP::twiddle(a.hi[e >> log_b], a.lo[e & ((1u << log_b) - 1)], a.rt);
            }
        }
    }

    // ---- sub-step 1: global -> registers -> DFT_r1 -> twiddle -> LDS -------
    {
        const uint32_t c = tid & (C - 1), g = tid >> LOGC;
        T v[r1];
        const T *src = in + t0 + c + (uint64_t)g * stride_in;
#pragma unroll
        for (int m = 0; m < r1; ++m) v[m] = FP::load_in(src[(uint64_t)m * (R / r1) * stride_in], a.rt, a.flags);
        const auto roots = FP::template load_roots<LR1, DIR>(a.hi, a.log_rmax);
        dft_regs<FP, LR1, DIR>(v, roots, a.rt);
#pragma unroll
        for (int k = 0; k < r1; ++k) {
            T x = v[bitrev_c(k, LR1)];
            // hi[0] = 1: lanes with g = 0 multiply by one rather than diverge
            if (k != 0) x = FP::mul(x, a.hi[(g * k) << (a.log_rmax - LOGR)], a.rt);
            const uint32_t row = r1 * g + k;
            lds[row * C + (c ^ (row & (C - 1)))] = x;
        }
    }
    __syncthreads();

    // ---- sub-step 2: LDS -> registers -> DFT_r2 -> pass twiddle -> global --
    const auto roots2 = FP::template load_roots<LR2, DIR>(a.hi, a.log_rmax);
#pragma unroll
    for (int gi = 0; gi < G2; ++gi) {
        const uint32_t q = tid + gi * NT;
        uint32_t c, g2;
        if constexpr (J1) {  // outputs of a column are contiguous: lanes walk g2
            g2 = q & (r1 - 1);
            c = q >> LR1;
        } else {  // outputs of a row are contiguous: lanes walk c
            c = q & (C - 1);
            g2 = q >> LOGC;
        }
        T v[r2];
#pragma unroll
        for (int m = 0; m < r2; ++m) {
            const uint32_t row = g2 + m * r1;
            v[m] = lds[row * C + (c ^ (row & (C - 1)))];
        }
        dft_regs<FP, LR2, DIR>(v, roots2, a.rt);
#pragma unroll
        for (int k2 = 0; k2 < r2; ++k2) {
            const uint32_t k = k2 * r1 + g2;
            T x = v[bitrev_c(k2, LR2)];
            // uniform code: hi[0] = lo[0] = twl[0] = 1
            if constexpr (J1) {
                const uint32_t e = (t0 + c) * k;  // s = t0 + c, J = 1
                const uint32_t log_b = logn - a.log_rmax;
                x = FP::mul(x, a.hi[e >> log_b], a.rt);
                if (log_b) x = FP::mul(x, a.lo[e & ((1u << log_b) - 1)], a.rt);
            } else {
                if (sJ != 0) x = FP::mul(x, twl[k], a.rt);
            }
            if (a.flags & PASS_LAST) x = FP::finish(x, a.rt, a.flags);
            if constexpr (J1) out[(uint64_t)(t0 + c) * R + k] = x;
            else out[(uint64_t)(sJ * R) + ((uint64_t)k << a.log_j) + (t0 - sJ) + c] = x;
        }
    }
}

}  // namespace stntt
