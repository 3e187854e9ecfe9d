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
    return sizeof(typename FP::T) == 8 ? 4 : 5;  // 128-byte rows
}

template <int LOGR>
constexpr int tile_lr1() {
    return (LOGR + 1) / 2;
}

template <class FP, int LOGR>
constexpr int tile_threads() {
    return (1 << (LOGR - tile_lr1<LOGR>())) << tile_logc<FP>();  // r2 * C
}

// One tile of a pass, computed by NT cooperating threads (NT need not equal
// the natural (R/r1) C: threads loop over groups and idle ones predicate
// off).  `block` is the tile's linear index over the whole batch.  lds holds
// R * C words, twl R twiddles.
// J1: first pass of a schedule (J = 1): outputs of a column are contiguous
// (y[R t + k]); otherwise J >= C and a row of C outputs is contiguous.
//
// PERM_IN (first pass only): the input is in bit-reversed order.  Word
// q = t + m n/R of the transform sits at bitrev_L(q) = bitrev(t) R +
// bitrev_R(m), so column t is the contiguous R-word chunk at bitrev(t) R with
// its rows permuted; the tile is staged through LDS so global reads stay
// whole cache lines.  A pointwise operand (mul_in) is applied there too.
// PERM_OUT (last pass only, J R = n): the output goes to bit-reversed
// positions: DFT output k of column t lands at bitrev(t) R + bitrev_R(k), a
// contiguous chunk per column, written back through LDS.
enum : int { PERM_NONE = 0, PERM_IN = 1, PERM_OUT = 2 };

// Exchange halves.  A radix-512 tile is 64 KiB of LDS, which holds a CU to
// two workgroups.  For odd LOGR, r1 = 2 r2 and sub-step 2 has two groups per
// thread anyway, so the exchange can run in two rounds through a 32 KiB
// image: sub-step 1 keeps its outputs in registers, half of them go to LDS,
// every thread runs one sub-step 2 group of that half, then the other half
// follows.  Two more barriers, and twice the resident workgroups.  The
// split follows the stores: the first pass (whose threads write a column's
// outputs contiguously) splits the columns; later passes (whose threads
// write rows of C contiguous words) split the sub-step 2 groups g2, i.e. the
// rows, so each round still writes whole 128-byte rows.  Splitting their
// columns instead left every output row as two 64-byte halves written a
// round apart, and 13-22 % of the lines reached HBM twice.  Not for the
// permuted variants, which stage the whole tile.
template <class FP, int LOGR, int PERM>
constexpr bool tile_half() {
    constexpr int R = 1 << LOGR, C = 1 << tile_logc<FP>();
    return PERM == PERM_NONE && (size_t)R * C * sizeof(typename FP::T) > 32768 && (LOGR & 1);
}

// Waves per SIMD the register allocator must leave room for (1: no bound).
// Row halves hold half of sub-step 1's outputs through a sub-step 2 round;
// for the 64-bit Montgomery field that took radix 512 from 162 to 190
// registers (two waves per SIMD); bounded to three waves it fits in 168.
template <class FP, int LOGR, bool J1, int PERM>
constexpr int tile_min_waves() {
    return tile_half<FP, LOGR, PERM>() && !J1 ? 3 : 1;
}

// Inner twiddles from a per-tile LDS copy of the R roots w_R^e (exchange
// halves of fields without a cheap twiddle product): each thread reads its
// w_R^(g k) when it multiplies, instead of holding r1 - 1 gathered words
// from the start of the tile.
template <class FP, int LOGR, int PERM>
constexpr bool tile_inner_lds() {
    return tile_half<FP, LOGR, PERM>() && !FP::CHEAP_TWIDDLE_PRODUCT;
}

template <class FP, int LOGR, int DIR, bool J1, int NT, int PERM = PERM_NONE>
__device__ __forceinline__ void tile_body(const PassArgs<FP> &a, uint32_t block, typename FP::T *lds,
                                          typename FP::Tw *twl, typename FP::Tw *twi) {
    using T = typename FP::T;
    using Tw = typename FP::Tw;
    constexpr int LOGC = tile_logc<FP>();
    constexpr int C = 1 << LOGC;
    constexpr int R = 1 << LOGR;
    constexpr int LR1 = tile_lr1<LOGR>(), LR2 = LOGR - LR1;
    constexpr int r1 = 1 << LR1, r2 = 1 << LR2;
    constexpr int NG1 = (R / r1) * C, NG2 = r1 * C;  // groups of the two sub-steps
    constexpr int G1 = (NG1 + NT - 1) / NT, G2 = (NG2 + NT - 1) / NT;
    constexpr int TWL = (R + NT - 1) / NT;  // pass twiddles built per thread
    constexpr int SP = (R * C + NT - 1) / NT;  // staged words per thread (PERM_IN / PERM_OUT)
    static_assert(PERM != PERM_IN || (J1 && G1 == 1), "bit-reversed input: first pass, one group per thread");
    static_assert(PERM != PERM_OUT || !J1, "bit-reversed output: last pass only");
    constexpr bool HALF = tile_half<FP, LOGR, PERM>();
    static_assert(!HALF || (NT == tile_threads<FP, LOGR>() && G1 == 1 && NG2 == 2 * NT), "exchange halves layout");
    // halves by columns (first pass) or by rows (later passes)
    constexpr bool HALF_COLS = HALF && J1, HALF_ROWS = HALF && !J1;
    constexpr int CL = HALF_COLS ? C / 2 : C;  // columns of the LDS image
    constexpr int RS = HALF_ROWS ? r1 / 2 : r1;  // image rows between the m-th inputs of a sub-step 2 group
    constexpr bool INNER_LDS = tile_inner_lds<FP, LOGR, PERM>();
    static_assert(!INNER_LDS || (G1 == 1 && R % NT == 0), "inner twiddle table layout");
    // LDS image index of (row, image column lc): XOR swizzle by the row.  The
    // column-halves image of 4-byte words (16 columns, 64-byte rows, two per
    // 32 banks) swizzles by row / 2 instead: by the row itself, sub-step 2
    // lanes g2 and g2 + 16 of a half-wave met on one bank (512 conflict
    // cycles per tile); by row / 2 every half-wave access is conflict-free
    // and the XOR operand is still a per-thread or compile-time constant.
    auto img = [](uint32_t row, uint32_t lc) -> uint32_t {
        if constexpr (HALF_COLS && sizeof(T) == 4) return row * CL + (lc ^ ((row >> 1) & (CL - 1)));
        else return row * CL + (lc ^ (row & (CL - 1)));
    };

    const uint32_t tid = threadIdx.x;
    const uint32_t logn = a.log_n;
    const uint32_t tpt = a.tiles_per_transform;
    const uint32_t b = block / tpt;
    const uint32_t t0 = (block % tpt) << LOGC;  // first column t of the tile
    const uint64_t base = (uint64_t)b << logn;
    const uint32_t stride_in = 1u << (logn - LOGR);  // n / R
    const uint32_t log_b = logn - a.log_rmax;        // w^e = hi[e >> log_b] lo[e & bmask]
    const uint32_t bmask = (1u << log_b) - 1;
    const T *in = a.in_tab ? a.in_tab[b] : a.in + base;
    T *out = a.out_tab ? a.out_tab[b] : a.out + base;
    // all columns of the tile share s, hence the pass twiddles w^(s J k)
    const uint32_t sJ = J1 ? 0u : ((t0 >> a.log_j) << a.log_j);

    // ---- sub-step 1: global -> registers -> DFT_r1 -> twiddle -> LDS -------
    // Every global load of a group is issued before any of them is used (the
    // data rows, the inner twiddles w_R^(g k) and, with the first group, the
    // factors of the pass twiddles), so a tile waits for memory once rather
    // than once per dependent load.  Row bases are uniform pointers; a lane
    // adds one 32-bit element offset.
    const uint64_t mstep = (uint64_t)(R / r1) * stride_in;
    // 32-bit byte offsets from uniform row bases for 4-byte words.  A lane's
    // offset within a row is below n / min(r1, r2) words (input lanes
    // c + g n/R with g < R/r1; output rows g2 J with g2 < r1 and J <= n/R):
    // < 2^30 bytes for n <= 2^30; twiddle table offsets are below 2^21 bytes
    // (hi: 2^14 words, lo: n / 2^14).  The 32-bit field's 2^24 passes run
    // 2-3 % faster with them; for 8-byte words at R >= 256 (where the
    // offsets stay below 2^31 bytes up to n = 2^32) the Goldilocks passes
    // measured about 1 % slower, so those keep 64-bit addresses.
    constexpr bool OFF32 = sizeof(T) == 4;
    const auto roots = FP::template load_roots<LR1, DIR>(a.hi, a.log_rmax);
    const uint32_t lt = logn - LOGR;  // bits of a column index t
    if constexpr (PERM == PERM_IN) {
        // static_for rather than an unroll pragma: with the 64-bit Montgomery
        // product the unrolled loop exceeds the compiler's size limit, the
        // loop stays rolled and sv/su move to scratch memory
        T sv[SP], su[SP];
        const bool pw = a.mul_in != nullptr;
        const T *mul = pw ? a.mul_in + base : nullptr;
        static_for<SP>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const uint32_t idx = tid + i * NT;
            if ((R * C) % NT == 0 || idx < (uint32_t)(R * C)) {
                const uint32_t p = idx & (R - 1), c = idx >> LOGR;
                const uint32_t off = brev_bits(t0 + c, lt) * R + p;
                sv[i] = at_off(in, off);
                if (pw) su[i] = at_off(mul, off);
            }
        });
        __builtin_amdgcn_sched_barrier(0);
        static_for<SP>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const uint32_t idx = tid + i * NT;
            if ((R * C) % NT == 0 || idx < (uint32_t)(R * C)) {
                const uint32_t p = idx & (R - 1), c = idx >> LOGR;
                T x = FP::load_in(sv[i], a.rt, a.flags);
                if (pw) x = FP::mulg(x, su[i], a.rt);
                const uint32_t m = brev_bits(p, LOGR);
                lds[m * C + (c ^ (m & (C - 1)))] = x;
            }
        });
        __syncthreads();
    }
    [[maybe_unused]] T xo[HALF ? r1 : 1];  // HALF: sub-step 1 outputs until the thread's round
    [[maybe_unused]] uint32_t xo_c = 0, xo_g = 0;
#pragma unroll
    for (int gi = 0; gi < G1; ++gi) {
        const uint32_t q = tid + gi * NT;
        const bool act = NG1 % NT == 0 || q < (uint32_t)NG1;
        const uint32_t c = q & (C - 1), g = q >> LOGC;
        [[maybe_unused]] const uint32_t lane = t0 + c + g * stride_in;
        const uint32_t tstep = g << (a.log_rmax - LOGR);  // w_R^(g k) = hi[k tstep]
        T v[r1];
        Tw w1[r1];
        if (act) {
            if constexpr (PERM == PERM_IN) {
#pragma unroll
                for (int m = 0; m < r1; ++m) {
                    const uint32_t row = g + m * (R / r1);
                    v[m] = lds[row * C + (c ^ (row & (C - 1)))];
                }
            } else {
#pragma unroll
                for (int m = 0; m < r1; ++m) {
                    if constexpr (OFF32)
                        v[m] = at_b32(in + t0 + (uint64_t)m * mstep, (c + g * stride_in) * (uint32_t)sizeof(T));
                    else
                        v[m] = at_off(in + (uint64_t)m * mstep, lane);
                }
            }
            if constexpr (FP::CHEAP_TWIDDLE_PRODUCT) {
                w1[1] = at_off(a.hi, tstep);  // the rest follow as its powers
            } else if constexpr (!INNER_LDS) {
#pragma unroll
                for (int k = 1; k < r1; ++k) {
                    if constexpr (OFF32) w1[k] = at_b32(a.hi, tstep * k * (uint32_t)sizeof(Tw));
                    else w1[k] = at_off(a.hi, tstep * k);
                }
            }
        }
        [[maybe_unused]] Tw ti[INNER_LDS ? R / NT : 1];
        if constexpr (INNER_LDS) {
#pragma unroll
            for (int i = 0; i < R / NT; ++i) ti[i] = at_off(a.hi, (tid + i * NT) << (a.log_rmax - LOGR));
        }
        // keep the scheduler from pulling a dependent use (and its wait)
        // in between the loads
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (!J1) {
            if (gi == 0 && sJ != 0) {
                Tw fh[TWL], fl[TWL];
#pragma unroll
                for (int i = 0; i < TWL; ++i) {
                    const uint32_t k = tid + i * NT, e = sJ * k;  // e < n
                    if (R % NT == 0 || k < (uint32_t)R) {
                        fh[i] = at_off(a.hi, e >> log_b);
                        fl[i] = at_off(a.lo, e & bmask);
                    }
                }
#pragma unroll
                for (int i = 0; i < TWL; ++i) {
                    const uint32_t k = tid + i * NT;
                    if (R % NT == 0 || k < (uint32_t)R) twl[k] = FP::twiddle(fh[i], fl[i], a.rt);
                }
            }
        }
        if constexpr (INNER_LDS) {
#pragma unroll
            for (int i = 0; i < R / NT; ++i) twi[tid + i * NT] = ti[i];
        }
        if constexpr (PERM == PERM_IN) __syncthreads();  // staged rows all read before reuse
        if (act) {
            if constexpr (PERM != PERM_IN) {
#pragma unroll
                for (int m = 0; m < r1; ++m) v[m] = FP::load_in(v[m], a.rt, a.flags);
            }
            dft_regs<FP, LR1, DIR>(v, roots, a.rt);
        }
        if constexpr (INNER_LDS) __syncthreads();  // the root table is complete
        if (act) {
            [[maybe_unused]] Tw wk = w1[1];  // w_R^(g k), walked along k
#pragma unroll
            for (int k = 0; k < r1; ++k) {
                T x = v[bitrev_c(k, LR1)];
                if constexpr (FP::CHEAP_TWIDDLE_PRODUCT) {
                    if (k >= 2) wk = FP::twiddle(wk, w1[1], a.rt);
                    if (k != 0) x = FP::mul(x, wk, a.rt);
                } else if constexpr (INNER_LDS) {
                    if (k != 0) x = FP::mul(x, twi[g * k], a.rt);  // g k < R
                } else {
                    if (k != 0) x = FP::mul(x, w1[k], a.rt);
                }
                if constexpr (HALF) {
                    // row halves: computed here, not next to round 1's LDS
                    // write; sunk past the barrier, each output kept its
                    // operand and its twiddle live (156 -> 264 registers for
                    // Goldilocks radix 512)
                    if constexpr (HALF_ROWS) asm volatile("" : "+v"(x));
                    xo[k] = x;
                } else {
                    const uint32_t row = r1 * g + k;
                    lds[row * C + (c ^ (row & (C - 1)))] = x;
                }
            }
        }
        xo_c = c;
        xo_g = g;
    }
    if constexpr (!HALF) __syncthreads();

    // ---- sub-step 2: LDS -> registers -> DFT_r2 -> pass twiddle -> global --
    const auto roots2 = FP::template load_roots<LR2, DIR>(a.hi, a.log_rmax);
    // PERM_OUT: outputs are kept in registers until the LDS tile is free
    [[maybe_unused]] T xs[PERM == PERM_OUT ? G2 : 1][PERM == PERM_OUT ? r2 : 1];
    [[maybe_unused]] uint32_t xc[G2], xg2[G2];
    // one sub-step 2 group: column c, group g2; within the LDS image column
    // lc, group g2l (input m at image row g2l + m RS)
    auto substep2 = [&](uint32_t c, uint32_t g2, uint32_t lc, uint32_t g2l, int gi) {
        // J = 1: output k = k2 r1 + g2 of column s = t0 + c takes w^(s k),
        // walked as the running product w^(s g2) (w^(s r1))^k2.  Twiddle
        // products are exact in every field (Goldilocks words, Montgomery
        // words for the others), so these are the words a lookup would give,
        // from four loads per thread.  They are issued before the LDS reads so
        // their latency overlaps them and the DFT.
        [[maybe_unused]] Tw th[2], tl[2];
        if constexpr (J1) {
            const uint32_t s = t0 + c;
            const uint32_t e0 = s * g2, es = s * r1;  // < n
            th[0] = at_off(a.hi, e0 >> log_b);
            tl[0] = at_off(a.lo, e0 & bmask);
            th[1] = at_off(a.hi, es >> log_b);
            tl[1] = at_off(a.lo, es & bmask);
        }
        T v[r2];
#pragma unroll
        for (int m = 0; m < r2; ++m) {
            const uint32_t row = g2l + m * RS;
            v[m] = lds[img(row, lc)];
        }
        dft_regs<FP, LR2, DIR>(v, roots2, a.rt);
        [[maybe_unused]] Tw cur, step;
        if constexpr (J1) {
            cur = FP::twiddle(th[0], tl[0], a.rt);
            step = FP::twiddle(th[1], tl[1], a.rt);
            if constexpr (FP::SCALE_IN) {
                if (a.flags & PASS_SCALE_IN) cur = FP::twiddle(cur, FP::tw_ninv(a.rt), a.rt);  // uniform
            }
        }
        // stores: a uniform base per k2 plus one 32-bit lane offset
        const uint32_t lane_out = J1 ? (c * R + g2) : ((g2 << a.log_j) + c);
#pragma unroll
        for (int k2 = 0; k2 < r2; ++k2) {
            const uint32_t k = k2 * r1 + g2;
            T x = v[bitrev_c(k2, LR2)];
            if constexpr (J1) {
                x = FP::mul(x, cur, a.rt);
                if (k2 + 1 < r2) cur = FP::twiddle(cur, step, a.rt);
            } else {
                if (sJ != 0) x = FP::mul(x, twl[k], a.rt);  // uniform branch
            }
            if (a.flags & PASS_LAST) x = FP::finish(x, a.rt, a.flags);
            if constexpr (PERM == PERM_OUT) {
                xs[gi][k2] = x;
            } else {
                T *orow = J1 ? out + (uint64_t)t0 * R + k2 * r1
                             : out + (uint64_t)sJ * R + (t0 - sJ) + ((uint64_t)(k2 * r1) << a.log_j);
                if constexpr (OFF32) at_b32(orow, lane_out * (uint32_t)sizeof(T)) = x;
                else at_off(orow, lane_out) = x;
            }
        }
        xc[gi] = c;
        xg2[gi] = g2;
    };
    // the rounds and groups below expand at compile time (static_for): a
    // rolled loop would index xc / xg2 at run time (scratch memory) and, for
    // the 64-bit Montgomery field, the unroll pragma alone exceeds the
    // compiler's size limit
    if constexpr (HALF_COLS) {
        static_for<2>([&](auto hc) {
            constexpr int h = decltype(hc)::value;
            if ((xo_c >> (LOGC - 1)) == (uint32_t)h) {  // this thread's column is in half h
#pragma unroll
                for (int k = 0; k < r1; ++k) {
                    const uint32_t row = r1 * xo_g + k, lc = xo_c & (CL - 1);
                    lds[img(row, lc)] = xo[k];
                }
            }
            __syncthreads();
            // lanes walk g2: a column's outputs are contiguous
            const uint32_t g2 = tid & (r1 - 1), lc = tid >> LR1;
            substep2(h * CL + lc, g2, lc, g2, h);
            if constexpr (h == 0) __syncthreads();  // round 0's reads before round 1's writes
        });
    } else if constexpr (HALF_ROWS) {
        // round h: groups g2 in [h r1/2, (h+1) r1/2), i.e. sub-step 1 outputs
        // k in that range; output k of group g sits at image row g r1/2 + k'
        static_for<2>([&](auto hc) {
            constexpr int h = decltype(hc)::value;
#pragma unroll
            for (int k = 0; k < r1 / 2; ++k) {
                const uint32_t row = (r1 / 2) * xo_g + k;
                lds[row * C + (xo_c ^ (row & (C - 1)))] = xo[h * (r1 / 2) + k];
            }
            __syncthreads();
            // lanes walk c: a row of C outputs is contiguous
            const uint32_t c = tid & (C - 1), g2l = tid >> LOGC;
            substep2(c, h * (r1 / 2) + g2l, c, g2l, h);
            if constexpr (h == 0) __syncthreads();  // round 0's reads before round 1's writes
        });
    } else {
        static_for<G2>([&](auto gc) {
            constexpr int gi = decltype(gc)::value;
            const uint32_t q = tid + gi * NT;
            if (!(NG2 % NT == 0 || q < (uint32_t)NG2)) return;
            if constexpr (J1) {  // outputs of a column are contiguous: lanes walk g2
                substep2(q >> LR1, q & (r1 - 1), q >> LR1, q & (r1 - 1), gi);
            } else {  // outputs of a row are contiguous: lanes walk c
                substep2(q & (C - 1), q >> LOGC, q & (C - 1), q >> LOGC, gi);
            }
        });
    }
    if constexpr (PERM == PERM_OUT) {
        __syncthreads();  // every sub-step 2 read of the tile is done
#pragma unroll
        for (int gi = 0; gi < G2; ++gi) {
            if (NG2 % NT == 0 || tid + gi * NT < (uint32_t)NG2) {
#pragma unroll
                for (int k2 = 0; k2 < r2; ++k2) {
                    const uint32_t p = brev_bits(k2 * r1 + xg2[gi], LOGR);
                    lds[p * C + (xc[gi] ^ (p & (C - 1)))] = xs[gi][k2];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SP; ++i) {
            const uint32_t idx = tid + i * NT;
            if ((R * C) % NT == 0 || idx < (uint32_t)(R * C)) {
                const uint32_t p = idx & (R - 1), c = idx >> LOGR;
                at_off(out, brev_bits(t0 + c, lt) * R + p) = lds[p * C + (c ^ (p & (C - 1)))];
            }
        }
    }
}

template <class FP, int LOGR, int DIR, bool J1, int PERM = PERM_NONE>
__global__ void __launch_bounds__((tile_threads<FP, LOGR>()), (tile_min_waves<FP, LOGR, J1, PERM>()))
    ntt_tile_kernel(PassArgs<FP> a) {
    constexpr int C = 1 << tile_logc<FP>();
    constexpr int R = 1 << LOGR;
    __shared__ typename FP::T lds[R * C / (tile_half<FP, LOGR, PERM>() ? 2 : 1)];
    __shared__ typename FP::Tw twl[J1 ? 1 : R];
    __shared__ typename FP::Tw twi[tile_inner_lds<FP, LOGR, PERM>() ? R : 1];
    tile_body<FP, LOGR, DIR, J1, tile_threads<FP, LOGR>(), PERM>(a, blockIdx.x, lds, twl, twi);
}

}  // namespace stntt
