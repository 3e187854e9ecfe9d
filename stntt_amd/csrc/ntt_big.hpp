// ntt_big.hpp -- the large-radix pass: a radix-R Stockham step (R = 2^9 ..
// 2^12; the schedules use 2^10 .. 2^12) over tiles of up to 2^14 words, run
// by persistent workgroups that walk the tiles in an XCD-aware order.
//
// The pass applies (stage algebra in ntt_pass.hpp, restating
// src/ntt/ntt_cpu.cpp:66-79)
//     y[(R s + k) J + j] = w^(s J k) * sum_m x[s J + j + m n/R] w_R^(m k)
// to a tile of C consecutive columns t = s J + j and all R rows, R C <=
// 2^14 words (128 KiB of Goldilocks words, 64 KiB of 32-bit words), so 2^24
// points take two passes of radix 4096 instead of three of radix 256, and
// narrower tiles (C / 2, C / 4) give more workgroups for smaller n.  The
// R-point DFT of a column is
// three register sub-steps of radix 2^l1, 2^l2, 2^l3 (l1 = 4) with two LDS
// exchanges; sub-step 1 loads its rows straight from HBM and sub-step 3
// stores straight to HBM:
//   sub-step i, inner stride J_i (1, 2^l1, 2^(l1+l2)), radix r_i: group g <
//   R/r_i of a column reads rows g + m R/r_i (m < r_i), and output k goes to
//   row (r_i s + k) J_i + j (s = g / J_i, j = g mod J_i) times w_R^(s J_i k)
//   (none on sub-step 3, where s = 0), the in-tile Stockham order.
// Row segments are C w bytes (32 for Goldilocks and the 32-bit field at R =
// 4096): a 128-byte line is shared by 128 / (C w) neighbouring tiles, and
// those are placed on one XCD (workgroup i runs on XCD i mod 8) at the same
// time so the line is fetched from HBM once and read from that XCD's L2 by
// the others (copy probe, 2^24 8-byte words: 5.7 TB/s for the first pass's
// access shape and 5.0 TB/s for the last pass's, against 2.4 / 2.0 TB/s in
// plain block order and 6.6 TB/s for a plain copy).
//
// LDS: the tile, row-major (a row is C words), and a table of the R roots
// w_R^e for the inner twiddles, read where they are multiplied.  Sub-step 1
// writes rows 16 apart from neighbouring lanes (the transpose of the tile),
// which conflicts in the banks; every other access walks consecutive rows.
// An XOR swizzle of the rows removes the conflicts but its address
// arithmetic took ~30 more registers per lane, and the 1024-thread
// workgroup has 128: the conflicted writes cost far less than the VALU work
// they sit beside.  One workgroup of 1024 threads per CU (16 words each:
// 124 registers for Goldilocks); it walks tiles i, i + G, i + 2G, ... (G =
// grid size, a multiple of 8 so a workgroup stays on its XCD) and fills the
// root table once.  Loads and arithmetic overlap across CUs, which drift
// out of phase after the first tile; issuing the next tile's loads ahead
// (in registers) needs 32 more registers than a lane has.
#pragma once

#include "ntt_pass.hpp"

namespace stntt {

constexpr int BIG_MIN_LOGR = 9;
constexpr int BIG_MAX_LOGR = 12;
constexpr int BIG_TILE_BYTES = 128 * 1024;

// words per thread (one radix-16 group): 1024 threads hold 2^14 words, a
// 128 KiB tile of 8-byte words or a 64 KiB tile of 4-byte words
template <class FP>
constexpr int big_ept() {
    return 16;
}

// log2 of the tile columns for a full tile of 2^14 words
template <class FP>
constexpr int big_full_logc(int logr) {
    return 14 - logr;
}

// sub-step radices (log2): 12 -> 4 4 4, 11 -> 4 4 3, 10 -> 4 3 3, 9 -> 3 3 3
constexpr int big_sub_log(int logr, int i) {
    const int l1 = logr >= 10 ? 4 : 3;
    const int rest = logr - l1;
    return i == 0 ? l1 : (i == 1 ? (rest + 1) / 2 : rest / 2);
}

template <class FP, int LOGR, int LOGC>
constexpr int big_threads() {
    return (1 << (LOGR + LOGC)) / big_ept<FP>();
}

// XCD-aware order of the tiles: tiles that share 128-byte lines (SIB of
// them, consecutive in t) run on the same XCD at the same time.  A
// bijection of [0, ntiles): the last partial group of 8 SIB tiles keeps its
// order.
__device__ __forceinline__ uint32_t big_tile_of(uint32_t i, uint32_t ntiles, uint32_t log_sib) {
    const uint32_t group = 8u << log_sib;
    const uint32_t full = ntiles - ntiles % group;
    if (i >= full) return i;
    const uint32_t xcd = i & 7, local = i >> 3;
    const uint32_t s = local & ((1u << log_sib) - 1), L = ((local >> log_sib) << 3) + xcd;
    return (L << log_sib) + s;
}

// LDS index of (row, column)
template <int LOGC>
__device__ __forceinline__ uint32_t big_lds(uint32_t row, uint32_t c) {
    return (row << LOGC) + c;
}

template <class FP, int LOGR, int LOGC, int DIR, bool J1>
__global__ void __launch_bounds__((big_threads<FP, LOGR, LOGC>()))
    ntt_big_kernel(PassArgs<FP> a, uint32_t ntiles) {
    using T = typename FP::T;
    using Tw = typename FP::Tw;
    constexpr int R = 1 << LOGR, C = 1 << LOGC;
    constexpr int NT = big_threads<FP, LOGR, LOGC>();
    constexpr int EPT = big_ept<FP>();
    constexpr int L1 = big_sub_log(LOGR, 0), L2 = big_sub_log(LOGR, 1), L3 = big_sub_log(LOGR, 2);
    static_assert(L1 + L2 + L3 == LOGR, "sub-step split");
    constexpr int r1 = 1 << L1, r2 = 1 << L2, r3 = 1 << L3;
    constexpr int G1 = EPT / r1, G2 = EPT / r2, G3 = EPT / r3;  // groups per thread
    static_assert(G1 * r1 == EPT && G2 * r2 == EPT && G3 * r3 == EPT, "groups per thread");
    static_assert(NT * EPT == R * C && NT <= 1024, "tile threads");
    constexpr int LOG_SIB = (7 - LOGC - (sizeof(T) == 8 ? 3 : 2)) > 0 ? (7 - LOGC - (sizeof(T) == 8 ? 3 : 2)) : 0;

    __shared__ T lds[R * C];
    __shared__ Tw twr[R];  // w_R^e

    const uint32_t tid = threadIdx.x;
    const uint32_t logn = a.log_n;
    const uint32_t tpt = a.tiles_per_transform;
    const uint32_t stride_in = 1u << (logn - LOGR);  // n / R
    const uint32_t log_b = logn - a.log_rmax;        // w^e = hi[e >> log_b] lo[e & bmask]
    const uint32_t bmask = (1u << log_b) - 1;

    // the R roots for the inner twiddles, once per workgroup
    for (uint32_t e = tid; e < (uint32_t)R; e += NT) twr[e] = at_off(a.hi, e << (a.log_rmax - LOGR));
    const auto roots1 = FP::template load_roots<L1, DIR>(a.hi, a.log_rmax);
    const auto roots2 = FP::template load_roots<L2, DIR>(a.hi, a.log_rmax);
    const auto roots3 = FP::template load_roots<L3, DIR>(a.hi, a.log_rmax);

    // sub-step 1 loads of tile `T` into v[G1][r1]
    auto load_tile = [&](uint32_t tile, T (&v)[G1][r1]) {
        const uint32_t b = tile / tpt;
        const uint32_t t0 = (tile % tpt) << LOGC;
        const T *in = a.in_tab ? a.in_tab[b] : a.in + ((uint64_t)b << logn);
#pragma unroll
        for (int gi = 0; gi < G1; ++gi) {
            const uint32_t q = tid + gi * NT;
            const uint32_t c = q & (C - 1), g = q >> LOGC;
            const uint32_t lane = t0 + c + g * stride_in;
#pragma unroll
            for (int m = 0; m < r1; ++m) v[gi][m] = at_off(in + (uint64_t)m * ((uint64_t)(R / r1) * stride_in), lane);
        }
    };

    __syncthreads();  // the root table is complete
    for (uint32_t i = blockIdx.x; i < ntiles; i += gridDim.x) {
        const uint32_t tile = big_tile_of(i, ntiles, LOG_SIB);
        T v[G1][r1];
        load_tile(tile, v);

        const uint32_t b = tile / tpt;
        const uint32_t t0 = (tile % tpt) << LOGC;
        T *out = a.out_tab ? a.out_tab[b] : a.out + ((uint64_t)b << logn);
        const uint32_t sJ = J1 ? 0u : ((t0 >> a.log_j) << a.log_j);  // shared by the tile's columns

        // ---- sub-step 1: registers -> DFT_r1 -> w_R^(g k) -> LDS ---------
#pragma unroll
        for (int gi = 0; gi < G1; ++gi) {
            const uint32_t q = tid + gi * NT;
            const uint32_t c = q & (C - 1), g = q >> LOGC;
#pragma unroll
            for (int m = 0; m < r1; ++m) v[gi][m] = FP::load_in(v[gi][m], a.rt, a.flags);
            dft_regs<FP, L1, DIR>(v[gi], roots1, a.rt);
#pragma unroll
            for (int k = 0; k < r1; ++k) {
                T x = v[gi][bitrev_c(k, L1)];
                if (k != 0) x = FP::mul(x, twr[g * k], a.rt);  // g k < R
                lds[big_lds<LOGC>(r1 * g + k, c)] = x;
            }
        }
        __syncthreads();

        // ---- sub-step 2: LDS -> DFT_r2 -> w_R^(s J2 k) -> LDS -------------
        {
            T u[G2][r2];
#pragma unroll
            for (int gi = 0; gi < G2; ++gi) {
                const uint32_t q = tid + gi * NT;
                const uint32_t c = q & (C - 1), g = q >> LOGC;
#pragma unroll
                for (int m = 0; m < r2; ++m) u[gi][m] = lds[big_lds<LOGC>(g + m * (R / r2), c)];
            }
            __syncthreads();
#pragma unroll
            for (int gi = 0; gi < G2; ++gi) {
                const uint32_t q = tid + gi * NT;
                const uint32_t c = q & (C - 1), g = q >> LOGC;
                const uint32_t s = g >> L1, j = g & (r1 - 1);  // J2 = r1
                dft_regs<FP, L2, DIR>(u[gi], roots2, a.rt);
#pragma unroll
                for (int k = 0; k < r2; ++k) {
                    T x = u[gi][bitrev_c(k, L2)];
                    const uint32_t e = (s * k) << L1;  // s J2 k < R
                    if (k != 0) x = FP::mul(x, twr[e], a.rt);
                    lds[big_lds<LOGC>(((r2 * s + k) << L1) + j, c)] = x;
                }
            }
            __syncthreads();
        }

        // ---- sub-step 3: LDS -> DFT_r3 -> pass twiddle -> HBM ---------------
        {
            T u[G3][r3];
#pragma unroll
            for (int gi = 0; gi < G3; ++gi) {
                const uint32_t q = tid + gi * NT;
                const uint32_t c = q & (C - 1), g = q >> LOGC;
#pragma unroll
                for (int m = 0; m < r3; ++m) u[gi][m] = lds[big_lds<LOGC>(g + m * (R / r3), c)];
            }
#pragma unroll
            for (int gi = 0; gi < G3; ++gi) {
                const uint32_t q = tid + gi * NT;
                const uint32_t c = q & (C - 1), g = q >> LOGC;  // output k3 lands on row k3 R/r3 + g
                // pass twiddle w^(e k_out), e = t (first pass) or s J (later
                // passes), walked as w^(e g) (w^(e R/r3))^k3
                const uint32_t e = J1 ? (t0 + c) : sJ;
                Tw cur{}, step{};
                const bool tw = J1 || sJ != 0;
                if (tw) {
                    const uint32_t e0 = e * g, es = e * (uint32_t)(R / r3);  // < n
                    cur = FP::twiddle(at_off(a.hi, e0 >> log_b), at_off(a.lo, e0 & bmask), a.rt);
                    step = FP::twiddle(at_off(a.hi, es >> log_b), at_off(a.lo, es & bmask), a.rt);
                    if constexpr (FP::SCALE_IN) {
                        if (J1 && (a.flags & PASS_SCALE_IN)) cur = FP::twiddle(cur, FP::tw_ninv(a.rt), a.rt);
                    }
                }
                dft_regs<FP, L3, DIR>(u[gi], roots3, a.rt);
                const uint32_t lane_out = J1 ? (c * R + g) : ((g << a.log_j) + c);
#pragma unroll
                for (int k3 = 0; k3 < r3; ++k3) {
                    T x = u[gi][bitrev_c(k3, L3)];
                    if (tw) {
                        x = FP::mul(x, cur, a.rt);
                        if (k3 + 1 < r3) cur = FP::twiddle(cur, step, a.rt);
                    }
                    if (a.flags & PASS_LAST) x = FP::finish(x, a.rt, a.flags);
                    T *orow = J1 ? out + (uint64_t)t0 * R + (uint64_t)k3 * (R / r3)
                                 : out + (uint64_t)sJ * R + (t0 - sJ) + ((uint64_t)(k3 * (R / r3)) << a.log_j);
                    at_off(orow, lane_out) = x;
                }
            }
        }
        __syncthreads();  // every sub-step 3 read of this tile before the next tile's writes
    }
}

}  // namespace stntt
