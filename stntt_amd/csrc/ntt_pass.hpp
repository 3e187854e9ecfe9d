// ntt_pass.hpp -- one Stockham radix-R pass over LDS-resident tiles.
//
// The reference computes the DIF Stockham one radix-2 stage at a time
// (src/ntt/ntt_cpu.cpp:66-79; on the GPU one launch and one HBM pass per
// stage, src/cuda/ntt/implementations/st_ntt_global_radix2.cuh:95-117).
// Stage i (J = 2^i), pair q < n/2, j = q mod J:
//     y[2q - j] = a + b,   y[2q - j + J] = w^(q - j) (a - b),  a = x[q], b = x[q + n/2].
// r consecutive radix-2 stages compose into one radix-r step:
//     a_m = x[sJ + j + m n/r]                     (m < r)
//     y[(r s + k) J + j] = w^(s J k) * sum_m a_m w_r^(m k)   (k < r),
// with s < n/(rJ), j < J.  Because the arithmetic is exact, any grouping of
// the stages gives the reference's output bit for bit.  A pass applies one
// radix-R step (R up to 2^13) to a tile of C "columns" (consecutive t = sJ+j,
// or C whole transforms when R = n), with the tile staged in LDS:
//   1. global -> LDS, coalesced along whichever of (c, m) is contiguous;
//   2. the R-point DFT of every column as radix-2^k sub-steps (k <= 4):
//      each thread pulls r values of a column from LDS into registers, runs
//      an r-point DIF butterfly network in registers, multiplies by the
//      sub-step twiddle and writes back in Stockham order;
//   3. LDS -> global with the pass twiddle w^(s J k), and on the last pass
//      the inverse's n^-1 and canonicalisation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "field_gl.hpp"
#include "field_m64.hpp"
#include "field_u32.hpp"

namespace stntt {

// ---------------------------------------------------------------------------
// compile-time helpers
// ---------------------------------------------------------------------------
template <int N, class F, int I = 0>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, F, I + 1>(static_cast<F &&>(f));
    }
}

constexpr int bitrev_c(int x, int bits) {
    int y = 0;
    for (int i = 0; i < bits; ++i) y = (y << 1) | ((x >> i) & 1);
    return y;
}

// Sub-step split of an R = 2^LOGR point DFT into ceil(LOGR/4) register
// radices, as even as possible, larger ones first.
constexpr int substeps_of(int logr) { return (logr + 3) / 4; }
constexpr int substep_log(int logr, int i) {
    int ns = substeps_of(logr);
    return logr / ns + (i < logr % ns ? 1 : 0);
}
constexpr int substep_logj(int logr, int i) {
    int s = 0;
    for (int k = 0; k < i; ++k) s += substep_log(logr, k);
    return s;
}

// ---------------------------------------------------------------------------
// field policies
// ---------------------------------------------------------------------------
enum : uint32_t {
    PASS_FIRST = 1u,   // reduce raw inputs (32-bit field)
    PASS_LAST = 2u,    // canonicalise outputs, apply n^-1 for the inverse
    PASS_SIGNED = 4u,  // 32-bit inputs are signed ints (reference int storage)
    PASS_SCALE = 8u,   // multiply by n^-1 (inverse), on the last pass's store
    // multiply by n^-1 (inverse) through the first pass's pass twiddle: a J =
    // 1 pass multiplies every output word by its running-product twiddle, so
    // starting that product at n^-1 scales the whole transform for one
    // product per thread instead of one per word (fields with SCALE_IN)
    PASS_SCALE_IN = 64u,
    PASS_BITREV_IN = 16u,   // input in bit-reversed order (single-pass kernel)
    PASS_BITREV_OUT = 32u,  // output in bit-reversed order (single-pass kernel)
};

// Global (device) memory.  A generic pointer the compiler cannot trace to a
// kernel argument -- one read from a pointer table, or chosen between a
// table entry and an argument -- compiles to flat_load / flat_store, which
// count against the LDS wait counter too: every LDS wait and barrier of the
// pass then also waited for the outstanding HBM stores.
template <class T>
using global_t = __attribute__((address_space(1))) T;
// LDS, for a shared array reached through a function parameter
template <class T>
using shared_t = __attribute__((address_space(3))) T;

// element `elem` (a 32-bit index: n <= 2^32) of a uniform base pointer into
// global memory (data buffers and twiddle tables, never LDS); the byte
// offset is formed in 64 bits (8-byte words past 2^29 need 33 bits)
template <class T>
__device__ __forceinline__ const global_t<T> &at_off(const T *base, uint32_t elem) {
    return *(const global_t<T> *)((const char *)base + (uint64_t)elem * sizeof(T));
}
template <class T>
__device__ __forceinline__ global_t<T> &at_off(T *base, uint32_t elem) {
    return *(global_t<T> *)((char *)base + (uint64_t)elem * sizeof(T));
}

// element at a 32-bit byte offset from a uniform base pointer: the base
// stays a scalar register pair and the offset one 32-bit register
// (global_load / global_store with saddr), where at_off's 64-bit offset
// takes a 64-bit vector address per access.  For offsets below 2^32 only.
template <class T>
__device__ __forceinline__ const global_t<T> &at_b32(const T *base, uint32_t byte_off) {
    return *(const global_t<T> *)((const char *)base + byte_off);
}
template <class T>
__device__ __forceinline__ global_t<T> &at_b32(T *base, uint32_t byte_off) {
    return *(global_t<T> *)((char *)base + byte_off);
}

// A word of a twiddle table through the constant address space (4): the
// tables are never written while a transform runs, so a load at a uniform
// address becomes a scalar load into SGPRs instead of a vector load that
// holds a VGPR in every lane (the in-register DFT roots: 24 VGPRs for a
// 32-bit radix-512 tile, 48 for the 64-bit field).
template <class T>
__device__ __forceinline__ T ld_table(const T *base, uint32_t elem) {
    return *(const __attribute__((address_space(4))) T *)((const char *)base + (uint64_t)elem * sizeof(T));
}

// bit reversal of the low `bits` bits of x, 1 <= bits <= 32
__device__ __forceinline__ uint32_t brev_bits(uint32_t x, uint32_t bits) { return __brev(x) >> (32 - bits); }

struct FieldGL {
    using T = uint64_t;
    using Tw = uint64_t;
    struct Rt {
        uint64_t ninv;
    };
    // runtime roots are not needed: omega_r^j is a power of two
    struct Roots {};
    // a general product is ~25 VALU: inner twiddles come from the table
    static constexpr bool CHEAP_TWIDDLE_PRODUCT = false;
    static __device__ __forceinline__ T add(T a, T b, const Rt &) { return gl_add(a, b); }
    static __device__ __forceinline__ T sub(T a, T b, const Rt &) { return gl_sub(a, b); }
    static __device__ __forceinline__ T mul(T a, Tw w, const Rt &) { return gl_mul(a, w); }
    // product of two field elements (pointwise spectra)
    static __device__ __forceinline__ T mulg(T a, T b, const Rt &) { return gl_mul(a, b); }
    template <int LR, int DIR>
    static __device__ __forceinline__ Roots load_roots(const Tw *, int) {
        return {};
    }
    // (x - y) * omega_r^J: a shift by a power of two with the sign folded
    // into the subtraction
    template <int LR, int DIR, int J>
    static __device__ __forceinline__ T sub_root_mul(T x, T y, const Roots &, const Rt &) {
        return gl_sub_mul_pow2<gl_root_shift<LR, DIR>(J)>(x, y);
    }
    // raw words enter the canonical range once, on the first pass
    static __device__ __forceinline__ T load_in(T x, const Rt &, uint32_t flags) {
        return (flags & PASS_FIRST) ? gl_canon(x) : x;
    }
    static __device__ __forceinline__ T finish(T x, const Rt &rt, uint32_t flags) {
        return (flags & PASS_SCALE) ? gl_mul(x, rt.ninv) : x;
    }
    // twiddle hi * lo from the two factor tables
    static __device__ __forceinline__ Tw twiddle(Tw hi, Tw lo, const Rt &) { return gl_mul(hi, lo); }
    static __device__ __forceinline__ Tw tw_one(const Rt &) { return 1; }
    // n^-1 as a twiddle word (PASS_SCALE_IN)
    static constexpr bool SCALE_IN = true;
    static __device__ __forceinline__ Tw tw_ninv(const Rt &rt) { return rt.ninv; }
};

struct FieldU32 {
    using T = uint32_t;
    using Tw = uint32_t;  // w 2^32 mod p (Montgomery form)
    struct Rt {
        uint32_t p;     // modulus, odd, < 2^31
        uint32_t pinv;  // -p^-1 mod 2^32
        uint32_t bar;   // floor(2^32 / p)
        uint32_t c32;   // 2^32 mod p: also the Montgomery form of 1
        uint32_t ninv;  // n^-1 in Montgomery form
        uint32_t pad;
        uint64_t mu;    // floor(2^64 / p), for products of two data words
    };
    struct Roots {
        Tw w[16];  // omega_r^j, j < r/2 (r <= 32)
    };
    // a twiddle product (5 VALU) is cheaper than holding a gathered table
    // word per power in a register
    static constexpr bool CHEAP_TWIDDLE_PRODUCT = true;

    static __device__ __forceinline__ T add(T a, T b, const Rt &rt) { return u32_add(a, b, rt.p); }
    static __device__ __forceinline__ T sub(T a, T b, const Rt &rt) { return u32_sub(a, b, rt.p); }
    // data word (normal form) times twiddle: normal form
    static __device__ __forceinline__ T mul(T a, Tw w, const Rt &rt) { return u32_mont(a, w, rt.p, rt.pinv); }
    static __device__ __forceinline__ T mulg(T a, T b, const Rt &rt) { return u32_mul_any(a, b, rt.p, rt.mu); }
    // omega_r^j = hi[j << (log_rmax - LR)] (hi[e] = omega_Rmax^e)
    template <int LR, int DIR>
    static __device__ __forceinline__ Roots load_roots(const Tw *hi, int log_rmax) {
        Roots r;
#pragma unroll
        for (int j = 0; j < (1 << LR) / 2; ++j) r.w[j] = ld_table(hi, (uint32_t)(j << (log_rmax - LR)));
        return r;
    }
    template <int LR, int DIR, int J>
    static __device__ __forceinline__ T sub_root_mul(T x, T y, const Roots &r, const Rt &rt) {
        if constexpr (J == 0) return u32_sub(x, y, rt.p);
        return u32_mont(u32_sub(x, y, rt.p), r.w[J], rt.p, rt.pinv);
    }
    static __device__ __forceinline__ T load_in(T x, const Rt &rt, uint32_t flags) {
        if (flags & PASS_FIRST) return u32_reduce_input(x, rt.p, rt.bar, rt.c32, (flags & PASS_SIGNED) != 0);
        return x;
    }
    static __device__ __forceinline__ T finish(T x, const Rt &rt, uint32_t flags) {
        if (flags & PASS_SCALE) x = u32_mont(x, rt.ninv, rt.p, rt.pinv);
        return x;
    }
    // twiddle hi * lo: the Montgomery product of two Montgomery words
    static __device__ __forceinline__ Tw twiddle(Tw hi, Tw lo, const Rt &rt) { return u32_mont(hi, lo, rt.p, rt.pinv); }
    static __device__ __forceinline__ Tw tw_one(const Rt &rt) { return rt.c32; }
    // n^-1 as a twiddle word (PASS_SCALE_IN): rt.ninv is a Montgomery word
    static constexpr bool SCALE_IN = true;
    static __device__ __forceinline__ Tw tw_ninv(const Rt &rt) { return rt.ninv; }
};

// The 32-bit field for moduli below 2^30 with lazy reduction (field_u32.hpp):
// the same twiddle words, tables and constants as FieldU32, data words in
// [0, 2p) inside a transform.  The first pass brings raw inputs to [0, p) as
// FieldU32 does; intermediate passes store [0, 2p) words; the last pass
// (PASS_LAST, finish) returns them to [0, p), so every output word is the
// one FieldU32 computes.  The launchers pick it for p < 2^30.
struct FieldU32L : FieldU32 {
    static __device__ __forceinline__ T add(T a, T b, const Rt &rt) { return u32l_add(a, b, 2 * rt.p); }
    static __device__ __forceinline__ T sub(T a, T b, const Rt &rt) { return u32l_sub(a, b, 2 * rt.p); }
    static __device__ __forceinline__ T mul(T a, Tw w, const Rt &rt) { return u32l_mont(a, w, rt.p, rt.pinv); }
    // u32_mul_any's Barrett quotient is exact to one for any product below
    // 2^64, so [0, 2p) operands give a [0, p) result
    static __device__ __forceinline__ T mulg(T a, T b, const Rt &rt) { return u32_mul_any(a, b, rt.p, rt.mu); }
    template <int LR, int DIR, int J>
    static __device__ __forceinline__ T sub_root_mul(T x, T y, const Roots &r, const Rt &rt) {
        if constexpr (J == 0) return u32l_sub(x, y, 2 * rt.p);
        return u32l_sub_mont(x, y, r.w[J], rt.p, rt.pinv);
    }
    static __device__ __forceinline__ T load_in(T x, const Rt &rt, uint32_t flags) {
        if (flags & PASS_FIRST) return u32_reduce_input(x, rt.p, rt.bar, rt.c32, (flags & PASS_SIGNED) != 0);
        return x;
    }
    static __device__ __forceinline__ T finish(T x, const Rt &rt, uint32_t flags) {
        if (flags & PASS_SCALE) x = u32l_mont(x, rt.ninv, rt.p, rt.pinv);
        return u32l_canon(x, rt.p);
    }
    // running products of twiddles: both factors below 2p, product below 2p
    static __device__ __forceinline__ Tw twiddle(Tw hi, Tw lo, const Rt &rt) { return u32l_mont(hi, lo, rt.p, rt.pinv); }
};

// Any odd prime below 2^64 (field_m64.hpp): data and twiddles in Montgomery
// form inside a transform, converted on the first load and the last store.
struct FieldM64 {
    using T = uint64_t;
    using Tw = uint64_t;  // w R mod p
    using Rt = M64Rt;
    struct Roots {
        Tw w[16];  // omega_r^j (Montgomery form), j < r/2 (r <= 32)
    };
    static constexpr bool CHEAP_TWIDDLE_PRODUCT = false;
    static __device__ __forceinline__ T add(T a, T b, const Rt &rt) { return m64_add(a, b, rt.p); }
    static __device__ __forceinline__ T sub(T a, T b, const Rt &rt) { return m64_sub(a, b, rt.p); }
    static __device__ __forceinline__ T mul(T a, Tw w, const Rt &rt) { return m64_mul(a, w, rt.p, rt.pinv); }
    // a (any form) times b (normal form), keeping a's form
    static __device__ __forceinline__ T mulg(T a, T b, const Rt &rt) {
        return m64_mul(a, m64_mul(b, rt.r2, rt.p, rt.pinv), rt.p, rt.pinv);
    }
    template <int LR, int DIR>
    static __device__ __forceinline__ Roots load_roots(const Tw *hi, int log_rmax) {
        Roots r;
#pragma unroll
        for (int j = 0; j < (1 << LR) / 2; ++j) r.w[j] = ld_table(hi, (uint32_t)(j << (log_rmax - LR)));
        return r;
    }
    template <int LR, int DIR, int J>
    static __device__ __forceinline__ T sub_root_mul(T x, T y, const Roots &r, const Rt &rt) {
        if constexpr (J == 0) return m64_sub(x, y, rt.p);
        return m64_mul(m64_sub(x, y, rt.p), r.w[J], rt.p, rt.pinv);
    }
    // any 64-bit word -> Montgomery form, exactly
    static __device__ __forceinline__ T load_in(T x, const Rt &rt, uint32_t flags) {
        return (flags & PASS_FIRST) ? m64_mul(x, rt.r2, rt.p, rt.pinv) : x;
    }
    // Montgomery form -> normal form, times n^-1 for the inverse
    static __device__ __forceinline__ T finish(T x, const Rt &rt, uint32_t flags) {
        return m64_mul(x, (flags & PASS_SCALE) ? rt.ninv : 1, rt.p, rt.pinv);
    }
    static __device__ __forceinline__ Tw twiddle(Tw hi, Tw lo, const Rt &rt) { return m64_mul(hi, lo, rt.p, rt.pinv); }
    static __device__ __forceinline__ Tw tw_one(const Rt &rt) { return rt.r1; }
    // the last pass's conversion out of Montgomery form is a product anyway
    // and takes n^-1 for free, so the scale stays there
    static constexpr bool SCALE_IN = false;
    static __device__ __forceinline__ Tw tw_ninv(const Rt &rt) { return rt.r1; }
};

// ---------------------------------------------------------------------------
// in-register r-point DFT: DIF radix-2 network, natural order in, output k
// found at v[bitrev(k)].
// ---------------------------------------------------------------------------
template <class FP, int LR, int DIR>
__device__ __forceinline__ void dft_regs(typename FP::T *v, const typename FP::Roots &roots,
                                         const typename FP::Rt &rt) {
    constexpr int r = 1 << LR;
    static_for<LR>([&](auto st) {
        constexpr int h = r >> (decltype(st)::value + 1);  // half block size
        static_for<r / 2>([&](auto bi) {
            constexpr int b = decltype(bi)::value;
            constexpr int i = b % h;                     // position in half block
            constexpr int base = (b / h) * 2 * h;
            constexpr int jroot = i * (r / (2 * h));     // exponent of omega_r
            auto x = v[base + i], y = v[base + i + h];
            v[base + i] = FP::add(x, y, rt);
            v[base + i + h] = FP::template sub_root_mul<LR, DIR, jroot>(x, y, roots, rt);
        });
    });
}

// ---------------------------------------------------------------------------
// pass arguments
// ---------------------------------------------------------------------------
template <class FP>
struct PassArgs {
    const typename FP::T *in;
    typename FP::T *out;
    const typename FP::Tw *hi;  // hi[e] = omega^((n / Rmax) e), e < Rmax
    const typename FP::Tw *lo;  // lo[e] = omega^e, e < n / Rmax
    uint32_t log_n;
    uint32_t log_j;     // J = product of the radices of earlier passes
    uint32_t log_rmax;  // Rmax = min(n, 4096)
    uint32_t batches;   // transforms in the buffer
    uint32_t tiles_per_transform;
    uint32_t flags;
    typename FP::Rt rt;
    // first pass only: when set, input word i is multiplied by mul_in[i]
    // before the transform (a pointwise product fused into the load)
    const typename FP::T *mul_in;
    // gathered batches: when set, transform b reads from in_tab[b] (first
    // pass) / writes to out_tab[b] (last pass) instead of in + b n / out + b n
    const typename FP::T *const *in_tab;
    typename FP::T *const *out_tab;
};

// w^e for e < n through the two factor tables
template <class FP>
__device__ __forceinline__ typename FP::T apply_twiddle(typename FP::T x, uint64_t e, const PassArgs<FP> &a) {
    uint32_t log_b = a.log_n - a.log_rmax;
    uint64_t e_hi = e >> log_b, e_lo = e & ((1ull << log_b) - 1);
    if (e_hi) x = FP::mul(x, a.hi[e_hi], a.rt);
    if (e_lo) x = FP::mul(x, a.lo[e_lo], a.rt);
    return x;
}

// One sub-step I of the in-LDS R-point DFT (then recurse to I + 1).
template <class FP, int LOGR, int LOGC, int NT, int DIR, int I>
__device__ __forceinline__ void lds_substeps(typename FP::T *lds, const PassArgs<FP> &a) {
    using T = typename FP::T;
    constexpr int R = 1 << LOGR, C = 1 << LOGC;
    constexpr int LR = substep_log(LOGR, I);
    constexpr int r = 1 << LR;
    constexpr int LOGJ = substep_logj(LOGR, I);  // inner J'
    constexpr bool LASTSUB = (LOGJ + LR == LOGR);
    constexpr int NG = (R / r) * C;              // radix-r groups in the tile
    constexpr int G = (NG + NT - 1) / NT;        // groups per thread
    constexpr int STRIDE = R / r;

    const auto roots = FP::template load_roots<LR, DIR>(a.hi, a.log_rmax);
    T v[G][r];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        int q = threadIdx.x + gi * NT;
        if (NG % NT == 0 || q < NG) {
            int c = q & (C - 1), g = q >> LOGC;
#pragma unroll
            for (int m = 0; m < r; ++m) v[gi][m] = lds[(g + m * STRIDE) * C + c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        int q = threadIdx.x + gi * NT;
        if (NG % NT == 0 || q < NG) {
            int c = q & (C - 1), g = q >> LOGC;
            int s = g >> LOGJ, j = g & ((1 << LOGJ) - 1);
            dft_regs<FP, LR, DIR>(v[gi], roots, a.rt);
#pragma unroll
            for (int k = 0; k < r; ++k) {
                T x = v[gi][bitrev_c(k, LR)];
                if constexpr (!LASTSUB) {
                    // omega_R^(s J' k) = hi[(s J' k) << (log_rmax - LOGR)]
                    uint32_t e = (uint32_t)((s << LOGJ) * k);
                    if (k != 0 && e != 0) x = FP::mul(x, a.hi[e << (a.log_rmax - LOGR)], a.rt);
                }
                lds[((r * s + k) * (1 << LOGJ) + j) * C + c] = x;
            }
        }
    }
    __syncthreads();
    if constexpr (!LASTSUB) lds_substeps<FP, LOGR, LOGC, NT, DIR, I + 1>(lds, a);
}

// Multi-pass tile: C consecutive t = sJ + j of one transform.
// Single-pass tile (R = n): C whole transforms of the batch.
template <class FP, int LOGR, int LOGC, int NT, int DIR, bool SINGLE>
__global__ void __launch_bounds__(NT) ntt_pass_kernel(PassArgs<FP> a) {
    using T = typename FP::T;
    constexpr int R = 1 << LOGR, C = 1 << LOGC;
    __shared__ T lds[R * C];

    const uint64_t n = 1ull << a.log_n;
    uint64_t in_base, out_base, cs_in, ms_in, cs_out, ks_out;
    uint32_t ncols = C;
    uint64_t t0 = 0;
    bool j_is_one = a.log_j == 0;
    if constexpr (SINGLE) {
        uint64_t b0 = (uint64_t)blockIdx.x * C;
        in_base = out_base = b0 << a.log_n;
        cs_in = cs_out = n;
        ms_in = ks_out = 1;
        uint64_t left = a.batches - b0;
        ncols = left < (uint64_t)C ? (uint32_t)left : (uint32_t)C;
    } else {
        uint64_t b = blockIdx.x / a.tiles_per_transform;
        uint64_t ti = blockIdx.x % a.tiles_per_transform;
        t0 = ti * C;
        in_base = (b << a.log_n) + t0;
        cs_in = 1;
        ms_in = n >> LOGR;
        if (j_is_one) {
            out_base = (b << a.log_n) + t0 * R;
            cs_out = R;
            ks_out = 1;
        } else {
            uint64_t s = t0 >> a.log_j, j0 = t0 & ((1ull << a.log_j) - 1);
            out_base = (b << a.log_n) + ((s * R) << a.log_j) + j0;
            cs_out = 1;
            ks_out = 1ull << a.log_j;
        }
    }

    // 1. global -> LDS
    const bool in_cfast = (cs_in == 1);
    for (int idx = threadIdx.x; idx < R * C; idx += NT) {
        int c, m;
        if (in_cfast) {
            c = idx & (C - 1);
            m = idx >> LOGC;
        } else {
            m = idx & (R - 1);
            c = idx >> LOGR;
        }
        T x = 0;
        if ((uint32_t)c < ncols) {
            const uint64_t gi = in_base + c * cs_in + m * ms_in;
            T raw;
            if constexpr (SINGLE) raw = a.in_tab ? at_off(a.in_tab[(uint64_t)blockIdx.x * C + c], (uint32_t)m) : a.in[gi];
            else raw = a.in[gi];
            x = FP::load_in(raw, a.rt, a.flags);
            if (a.mul_in) x = FP::mulg(x, a.mul_in[gi], a.rt);
        }
        // bit-reversed input: word m of a transform is DFT input bitrev(m)
        if constexpr (SINGLE) {
            if (a.flags & PASS_BITREV_IN) m = (int)brev_bits((uint32_t)m, LOGR);
        }
        lds[m * C + c] = x;
    }
    __syncthreads();

    // 2. R-point DFT of every column
    lds_substeps<FP, LOGR, LOGC, NT, DIR, 0>(lds, a);

    // 3. LDS -> global with the pass twiddle w^(s J k)
    const bool out_cfast = (cs_out == 1);
    for (int idx = threadIdx.x; idx < R * C; idx += NT) {
        int c, k;
        if (out_cfast) {
            c = idx & (C - 1);
            k = idx >> LOGC;
        } else {
            k = idx & (R - 1);
            c = idx >> LOGR;
        }
        if ((uint32_t)c >= ncols) continue;
        // bit-reversed output: word k of a transform is DFT output bitrev(k)
        int kk = k;
        if constexpr (SINGLE) {
            if (a.flags & PASS_BITREV_OUT) kk = (int)brev_bits((uint32_t)k, LOGR);
        }
        T x = lds[kk * C + c];
        if constexpr (!SINGLE) {
            // s J = t - j: with J = 1 it is t itself, otherwise the tile's s J
            uint64_t sJ = j_is_one ? (t0 + c) : ((t0 >> a.log_j) << a.log_j);
            uint64_t e = sJ * (uint64_t)k;
            if (e) x = apply_twiddle<FP>(x, e, a);
        }
        if (a.flags & PASS_LAST) x = FP::finish(x, a.rt, a.flags);
        if constexpr (SINGLE) {
            if (a.out_tab) {
                at_off(a.out_tab[(uint64_t)blockIdx.x * C + c], (uint32_t)k) = x;
                continue;
            }
        }
        a.out[out_base + c * cs_out + k * ks_out] = x;
    }
}

}  // namespace stntt
