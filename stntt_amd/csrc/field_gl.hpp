// field_gl.hpp -- Goldilocks field Z_P, P = 2^64 - 2^32 + 1, on gfx950 VALU.
//
// The reference declares this field (P and EPS at
// src/cuda/ntt/implementations/goldilocks/arithmetic.cuh:16-17, add/sub at
// :19-30) but leaves mul and the butterfly empty (:32-35).  This file builds
// the whole field from the special form of P:
//     2^64 == 2^32 - 1 = EPS (mod P),   2^96 == -1 (mod P).
// Values travel "weakly reduced": any 64-bit word is a valid representative
// of its class mod P.  Only gl_canon() maps to [0, P), once, at the final
// store.  Every routine is exact for all 64-bit inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stntt {

constexpr uint64_t GL_P = 0xFFFFFFFF00000001ull;
constexpr uint64_t GL_EPS = 0x00000000FFFFFFFFull;

// a + b (mod P).  A carry out of 2^64 is worth EPS; adding EPS can carry a
// second time only when a or b was >= P, and then a second EPS settles it.
__device__ __forceinline__ uint64_t gl_add(uint64_t a, uint64_t b) {
    uint64_t s;
    bool c1 = __builtin_add_overflow(a, b, &s);
    uint64_t t;
    bool c2 = __builtin_add_overflow(s, c1 ? GL_EPS : 0ull, &t);
    return t + (c2 ? GL_EPS : 0ull);
}

// a - b (mod P).  A borrow is worth -EPS; the second borrow case mirrors
// gl_add.
__device__ __forceinline__ uint64_t gl_sub(uint64_t a, uint64_t b) {
    uint64_t d;
    bool c1 = __builtin_sub_overflow(a, b, &d);
    uint64_t t;
    bool c2 = __builtin_sub_overflow(d, c1 ? GL_EPS : 0ull, &t);
    return t - (c2 ? GL_EPS : 0ull);
}

// lo + hi * 2^64 (mod P) with hi = h0 + h1 * 2^32:
//     == lo - h1 + h0 * EPS,   and  h0 * EPS = (h0 << 32) - h0  < 2^64.
__device__ __forceinline__ uint64_t gl_reduce128(uint64_t lo, uint64_t hi) {
    uint64_t h0 = hi & 0xFFFFFFFFull, h1 = hi >> 32;
    uint64_t t;
    bool b = __builtin_sub_overflow(lo, h1, &t);
    t -= b ? GL_EPS : 0ull;  // t was >= 2^64 - 2^32 + 1 > EPS: no second borrow
    uint64_t u = (h0 << 32) - h0;
    uint64_t r;
    bool c = __builtin_add_overflow(t, u, &r);
    return r + (c ? GL_EPS : 0ull);  // r < u here, so r + EPS < 2^64
}

__device__ __forceinline__ uint64_t gl_mul(uint64_t a, uint64_t b) {
    return gl_reduce128(a * b, __umul64hi(a, b));
}

// x * 2^S (mod P) for a compile-time S; 2^192 == 1 and 2^96 == -1, so only
// S mod 96 shifts and S mod 192 >= 96 negates.
template <int S>
__device__ __forceinline__ uint64_t gl_mul_pow2(uint64_t x) {
    constexpr int s192 = ((S % 192) + 192) % 192;
    constexpr bool neg = s192 >= 96;
    constexpr int s = s192 % 96;
    uint64_t r;
    if constexpr (s == 0) {
        r = x;
    } else if constexpr (s < 64) {
        r = gl_reduce128(x << s, x >> (64 - s));
    } else {
        // 2^s = 2^(s-32) * 2^32 with both shifts below 64
        uint64_t y = gl_reduce128(x << (s - 32), x >> (96 - s));
        r = gl_reduce128(y << 32, y >> 32);
    }
    if constexpr (neg) return gl_sub(0ull, r);
    return r;
}

// canonical representative in [0, P)
__device__ __forceinline__ uint64_t gl_canon(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

// omega_64 = 7^((P-1)/64) = 2^39 (mod P), so every root of unity of order
// r | 64 is a power of two: omega_r^j = 2^(j * 39 * 64 / r).  Inverse roots
// use the negated exponent.  Returns the exponent of 2 for omega_r^j.
template <int LOGR, int DIR>
constexpr int gl_root_shift(int j) {
    int e = (j * 39 * (64 >> LOGR)) % 192;
    return DIR == 0 ? e : (192 - e) % 192;
}

}  // namespace stntt
