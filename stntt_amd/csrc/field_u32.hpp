// field_u32.hpp -- the reference's 32-bit prime field on gfx950 VALU.
//
// The reference keeps values in signed int and reduces with C '%'
// (src/ntt/ntt_util.cpp:88-94, src/cuda/math.cuh:10-16); its products
// w * (a - b) overflow int32 from p = 65537 on (src/ntt/ntt_cpu.cpp:74,
// src/cuda/ntt/arithmetic.cuh:23).  Here p is any runtime odd prime below
// 2^31 (every findParams prime up to n = 2^27), values are canonical in
// [0, p), and products are exact:
//   * data words stay in normal form; twiddles are stored in Montgomery
//     form wm = w 2^32 mod p, so x * w = REDC(x wm) (Montgomery, "Modular
//     multiplication without trial division", Math. Comp. 44, 1985): one
//     32 x 32 -> 64 product, one mul_lo, one 64-bit multiply-add and a
//     conditional subtraction;
//   * the product of two twiddles, REDC(am bm) = (a b) 2^32 mod p, is again
//     a Montgomery twiddle, so twiddle runs are running products;
//   * a + b and a - b use the min(s, s - p) trick on unsigned words.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define U32_HD __host__ __device__ __forceinline__

namespace stntt {

U32_HD uint32_t u32_add(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t s = a + b;  // < 2p < 2^32
    uint32_t t = s - p;
    return s < t ? s : t;
}

U32_HD uint32_t u32_sub(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t d = a + p - b;  // in (0, 2p)
    uint32_t t = d - p;
    return d < t ? d : t;
}

// x wm 2^-32 mod p in [0, p) for any 32-bit x and wm < p (p odd, < 2^31),
// pinv = -p^-1 mod 2^32.  t = x wm < 2^63; m = t pinv mod 2^32 makes
// t + m p divisible by 2^32, t + m p < 2^64, and the quotient is below 2p.
U32_HD uint32_t u32_mont(uint32_t x, uint32_t wm, uint32_t p, uint32_t pinv) {
    const uint64_t t = (uint64_t)x * wm;
    const uint32_t m = (uint32_t)t * pinv;
    const uint32_t u = (uint32_t)(((uint64_t)m * p + t) >> 32);
    const uint32_t r = u - p;
    return u < r ? u : r;
}

// ---------------------------------------------------------------------------
// Lazy reduction for p < 2^30 (Harvey, "Faster arithmetic for number-
// theoretic transforms", J. Symbolic Comput. 60, 2014): words between
// operations lie in [0, 2p) instead of [0, p).  With 4p < 2^32 no sum or
// biased difference wraps, and a Montgomery product of x < 4p by wm < p (or
// of two words below 2p) is below 4p^2 < 2^32 p, so its quotient
// (t + m p) / 2^32 is already below 2p: the final subtraction is dropped.
// A butterfly's (a - b) w is one biased subtraction and a three-instruction
// product (5 VALU) instead of a reduced subtraction and a reduced product
// (9).  Words are brought to [0, p) once, where a transform stores its
// output.
// ---------------------------------------------------------------------------

// a + b for a, b < 2p: result in [0, 2p)
U32_HD uint32_t u32l_add(uint32_t a, uint32_t b, uint32_t p2) {
    uint32_t s = a + b;  // < 4p < 2^32
    uint32_t t = s - p2;
    return s < t ? s : t;
}

// a - b for a, b < 2p: result in [0, 2p)
U32_HD uint32_t u32l_sub(uint32_t a, uint32_t b, uint32_t p2) {
    uint32_t d = a + p2 - b;  // in (0, 4p)
    uint32_t t = d - p2;
    return d < t ? d : t;
}

// x wm 2^-32 mod p in [0, 2p) for x wm < 2^32 p
U32_HD uint32_t u32l_mont(uint32_t x, uint32_t wm, uint32_t p, uint32_t pinv) {
    const uint64_t t = (uint64_t)x * wm;
    const uint32_t m = (uint32_t)t * pinv;
    return (uint32_t)(((uint64_t)m * p + t) >> 32);
}

// (a - b) wm 2^-32 mod p in [0, 2p) for a, b < 2p and wm < p: the difference
// is biased by 2p into (0, 4p) and multiplied unreduced
U32_HD uint32_t u32l_sub_mont(uint32_t a, uint32_t b, uint32_t wm, uint32_t p, uint32_t pinv) {
    return u32l_mont(a + 2 * p - b, wm, p, pinv);
}

// [0, 2p) -> [0, p)
U32_HD uint32_t u32l_canon(uint32_t x, uint32_t p) {
    uint32_t t = x - p;
    return x < t ? x : t;
}

// the lazy form applies: 4p < 2^32
U32_HD bool u32_lazy_modulus(uint32_t p) { return p < (1u << 30); }

// -p^-1 mod 2^32 for odd p: Newton's iteration doubles the correct low bits
// (x = p is correct to 3 bits)
inline uint32_t u32_neg_inverse(uint32_t p) {
    uint32_t x = p;
    for (int i = 0; i < 4; ++i) x *= 2u - p * x;
    return 0u - x;
}

// w 2^32 mod p: the Montgomery form of a twiddle (host side)
inline uint32_t u32_to_mont(uint64_t w, uint32_t p) { return (uint32_t)(((w % p) << 32) % p); }

// high 64 bits of a 64 x 64 product
U32_HD uint64_t u64_mulhi(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// a * b mod p for two data words a, b < p: Barrett with mu = floor(2^64 / p).
// x = ab < 2^62; q = floor(x mu / 2^64) is floor(x / p) or one less, so
// x - q p lies in [0, 2p) and fits 32 bits.
U32_HD uint32_t u32_mul_any(uint32_t a, uint32_t b, uint32_t p, uint64_t mu) {
    uint64_t x = (uint64_t)a * b;
    uint64_t q = u64_mulhi(x, mu);
    uint32_t r = (uint32_t)x - (uint32_t)q * p;
    uint32_t t = r - p;
    return r < t ? r : t;
}

// Reduce an arbitrary 32-bit input to [0, p).  bar = floor(2^32 / p);
// x - floor(x * bar / 2^32) * p lies in [0, 2p).  For signed inputs a word
// with the top bit set stands for x - 2^32, i.e. minus (2^32 mod p) = c32;
// this reproduces the reference's modulo() on negative ints.
U32_HD uint32_t u32_reduce_input(uint32_t x, uint32_t p, uint32_t bar, uint32_t c32,
                                                     bool is_signed) {
    uint32_t r = x - (uint32_t)(((uint64_t)x * bar) >> 32) * p;
    uint32_t t = r - p;
    r = r < t ? r : t;
    if (is_signed && (int32_t)x < 0) r = u32_sub(r, c32, p);
    return r;
}

}  // namespace stntt
