// field_gl.hpp -- Goldilocks field Z_P, P = 2^64 - 2^32 + 1, on gfx950 VALU.
//
// The reference declares this field (P and EPS at
// src/cuda/ntt/implementations/goldilocks/arithmetic.cuh:16-17, add/sub at
// :19-30) but leaves mul and the butterfly empty (:32-35).  This file builds
// the whole field from the special form of P:
//     2^64 == 2^32 - 1 = EPS (mod P),   2^96 == -1 (mod P).
//
// Invariant: every value handed between routines is canonical, in [0, P).
// Raw inputs are brought there once (gl_canon) when a transform loads them.
// Keeping values canonical makes a + b and a - b one carry/borrow
// correction each (no double wrap is possible), which is cheaper overall
// than "any 64-bit word" representatives.
//
// On gfx950 the 32x32->64 multiply-add (v_mad_u64_u32), 64-bit add
// (v_lshl_add_u64) and 64-bit compare all issue at full VALU rate, so cost is
// simply instruction count; the sequences below are written for it.
// Everything is __host__ __device__ so the CPU tests can check it exhaustively
// on edge values.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GL_HD __host__ __device__ __forceinline__

namespace stntt {

constexpr uint64_t GL_P = 0xFFFFFFFF00000001ull;
constexpr uint64_t GL_EPS = 0x00000000FFFFFFFFull;

// x + EPS (mod 2^64) when c, else x: one 32-bit select and a 64-bit add of
// the zero-extended word, where selecting between x and x + EPS takes a
// select per half
GL_HD uint64_t gl_add_eps_if(uint64_t x, bool c) { return x + (uint64_t)(c ? 0xFFFFFFFFu : 0u); }

// any 64-bit word -> [0, P): x - P = x + EPS (mod 2^64) when x >= P
GL_HD uint64_t gl_canon(uint64_t x) { return gl_add_eps_if(x, x >= GL_P); }

// a + b, a, b < P.  a + b < 2P: subtract P when the 64-bit sum carried (the
// 2^64 is worth EPS, and s + EPS < P then) or reached P (s + EPS = s - P
// mod 2^64); both are s + EPS, and they never coincide.
GL_HD uint64_t gl_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    return gl_add_eps_if(s, (s < a) | (s >= GL_P));
}

// d = a - b (mod 2^64) as a two-word borrow chain; returns the borrow.  The
// chain's own borrow-out (VCC after v_subb_co_u32) selects the correction;
// a 64-bit __builtin_sub_overflow compiled to the same subtraction plus a
// 64-bit compare that recomputed it.
GL_HD bool gl_sub_borrow(uint64_t a, uint64_t b, uint64_t &d) {
    unsigned c1, c2;
    const uint32_t lo = __builtin_subc((uint32_t)a, (uint32_t)b, 0u, &c1);
    const uint32_t hi = __builtin_subc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c1, &c2);
    d = ((uint64_t)hi << 32) | lo;
    return c2 != 0;
}

// a - b, a, b < P.  A borrow means a - b + 2^64; adding P is subtracting
// EPS, which cannot borrow again (a - b + 2^64 > 2^32).
GL_HD uint64_t gl_sub(uint64_t a, uint64_t b) {
    uint64_t d;
    return gl_sub_borrow(a, b, d) ? d + GL_P : d;
}

// lo + hi * 2^64 (mod P) for any lo, hi; result in [0, P).
// hi = h0 + h1 2^32 and 2^64 == EPS, 2^96 == -1:  lo + h0 EPS - h1.
//   s = lo - h1                 a borrow is worth -EPS (s + P, no second
//                               borrow: lo - h1 + 2^64 > EPS)
//   r = h0 * EPS + s            one v_mad_u64_u32; s is any 64-bit word
// A carry of r is worth +EPS, and then r + EPS <= 2^64 - 2^32 - 1 < P is
// already canonical; without one, r - P = r + EPS (mod 2^64) is due exactly
// when r + EPS wraps.  Both corrections are the one word r + EPS (the other
// order -- multiply-add, subtract, canonicalise -- took four more VALU).
GL_HD uint64_t gl_reduce128(uint64_t lo, uint64_t hi) {
    const uint64_t h0 = (uint32_t)hi, h1 = hi >> 32;
    uint64_t s;
    if (gl_sub_borrow(lo, h1, s)) s -= GL_EPS;
    const uint64_t r = h0 * GL_EPS + s;
    return gl_add_eps_if(r, (r < s) | (r >= GL_P));
}

// 64 x 64 -> 128 from four 32 x 32 -> 64 multiply-adds
GL_HD void gl_mul_wide(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    uint64_t p00 = a0 * b0;
    uint64_t p01 = a0 * b1 + (p00 >> 32);
    uint64_t p10 = a1 * b0 + (uint32_t)p01;
    hi = a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (uint32_t)p00;
}

// a * b for any 64-bit a, b; result in [0, P)
GL_HD uint64_t gl_mul(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    gl_mul_wide(a, b, lo, hi);
    return gl_reduce128(lo, hi);
}

// x * 2^s (mod P) for 0 < s <= 32 and any x; result in [0, P).
// x << s = lo + h 2^64 with h < 2^s, so lo + h EPS by one multiply-add; after
// a carry r + EPS < 2^32 EPS < P is already canonical, and without one
// r - P = r + EPS (mod 2^64) is due exactly when r >= P, i.e. when r + EPS
// wraps: one correction either way.
GL_HD uint64_t gl_shl_small(uint64_t x, int s) {
    uint64_t lo = x << s;
    uint64_t h = x >> (64 - s);
    uint64_t r = h * GL_EPS + lo;
    return gl_add_eps_if(r, (r < lo) | (r >= GL_P));
}

// x * 2^S (mod P) for 32 < S < 64 and any x; result in [0, P).  With
// t = S - 32, x = xh 2^32 + xl and x 2^S = lo + hi 2^64: lo has a zero low
// word and high word lh = xl 2^t mod 2^32; hi = h0 + h1 2^32 with
// h0 = (x >> (32 - t)) mod 2^32 and h1 = xh >> (32 - t).  2^64 == 2^32 - 1
// and 2^96 == -1 give
//     x 2^S == lh 2^32 + h0 (2^32 - 1) - h1 = (lh + h0) 2^32 - (h0 + h1).
// The carry c of lh + h0 = a + c 2^32 is worth 2^64 == 2^32 - 1, so
//     x 2^S == (a + c) 2^32 - (h0 + h1 + c),
// with a + c < 2^32 (a <= 2^32 - 2 when c = 1) and h0 + h1 + c < 2^33.  The
// 64-bit difference is at most (2^32 - 1) 2^32 = P - 1 when it does not
// borrow, and a borrow (a negative difference above -2^33) adds P.  Three
// VALU fewer than the general 128-bit reduction of lo + hi 2^64.
template <int S>
GL_HD uint64_t gl_shl_mid(uint64_t x) {
    static_assert(S > 32 && S < 64, "shift range");
    constexpr int t = S - 32;
    const uint32_t xh = (uint32_t)(x >> 32);
    const uint32_t lh = (uint32_t)x << t;
    const uint32_t h0 = (uint32_t)(x >> (32 - t));
    const uint32_t h1 = xh >> (32 - t);
    unsigned c, cb;
    const uint32_t a = __builtin_addc(lh, h0, 0u, &c);
    const uint32_t b0 = __builtin_addc(h0, h1, c, &cb);
    uint64_t r;
    const bool borrow = gl_sub_borrow((uint64_t)(a + c) << 32, ((uint64_t)cb << 32) | b0, r);
    return borrow ? r + GL_P : r;
}

// x * 2^S (mod P) for a compile-time S and any x; result in [0, P).
// 2^192 == 1 and 2^96 == -1.
template <int S>
GL_HD uint64_t gl_mul_pow2(uint64_t x) {
    constexpr int s192 = ((S % 192) + 192) % 192;
    constexpr bool neg = s192 >= 96;
    constexpr int s = s192 % 96;
    if constexpr (s >= 64) {
        // 64 <= s < 96: x 2^s = -x 2^-v with v = 96 - s in (0, 32].  With
        // x = q 2^v + r (r < 2^v) and w = r 2^(32-v) < 2^32:
        //   x 2^-v = q + r 2^-v = q - r 2^(96-v) = q - w 2^64 == q - w EPS,
        // so x 2^s == w EPS - q: a shift, a multiply-add and one
        // subtraction (q <= 2^63 and w EPS < 2^64 - 2^33 are both < P).
        constexpr int v = 96 - s;
        const uint64_t q = x >> v;
        const uint64_t e = (uint64_t)((uint32_t)x << (32 - v)) * GL_EPS;
        return neg ? gl_sub(q, e) : gl_sub(e, q);
    } else {
        uint64_t r;
        if constexpr (s == 0) r = gl_canon(x);
        else if constexpr (s <= 32) r = gl_shl_small(x, s);
        else r = gl_shl_mid<s>(x);
        if constexpr (neg) return gl_sub(0, r);
        return r;
    }
}

// omega_64 = 7^((P-1)/64) = 2^39 (mod P), so every root of unity of order
// r | 64 is a power of two: omega_r^j = 2^(j * 39 * 64 / r).  Inverse roots
// use the negated exponent.  Returns the exponent of 2 for omega_r^j.
template <int LOGR, int DIR>
constexpr int gl_root_shift(int j) {
    int e = (j * 39 * (64 >> LOGR)) % 192;
    return DIR == 0 ? e : (192 - e) % 192;
}

// (a - b) * 2^S for a, b < P with the sign of 2^S (2^96 == -1) folded
// into the subtraction: (a - b)(-2^s) = (b - a) 2^s.  Result in [0, P).
template <int S>
GL_HD uint64_t gl_sub_mul_pow2(uint64_t a, uint64_t b) {
    constexpr int s192 = ((S % 192) + 192) % 192;
    if constexpr (s192 == 0) return gl_sub(a, b);
    else if constexpr (s192 == 96) return gl_sub(b, a);
    else if constexpr (s192 > 96) return gl_mul_pow2<s192 - 96>(gl_sub(b, a));
    else return gl_mul_pow2<s192>(gl_sub(a, b));
}

}  // namespace stntt
