// field_m64.hpp -- Z_p for any odd prime p < 2^64, in Montgomery form.
//
// The reference's parameter search (findParams, src/ntt/ntt_util.cpp:10-49)
// computes in size_t but its pow overflows above 2^32 (ntt_util.h:72-96), so
// it cannot produce or use 64-bit NTT primes.  This field takes any prime
// p = k 2^L + 1 below 2^64 with a runtime root: Montgomery multiplication
// with R = 2^64 (Montgomery, "Modular multiplication without trial
// division", Math. Comp. 44, 1985).
//
// Representation inside a transform: data words and twiddles are in
// Montgomery form (x R mod p); products stay in that form.  The first pass
// converts any 64-bit input word exactly (mont(x, R^2) = x R mod p holds for
// every x < 2^64, since x (R^2 mod p) < R p); the last pass converts back,
// folding n^-1 into the same product for the inverse.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_gl.hpp"  // gl_mul_wide: the plain 64 x 64 -> 128 product

#define M64_HD __host__ __device__ __forceinline__

namespace stntt {

struct M64Rt {
    uint64_t p;     // the prime, odd, < 2^64
    uint64_t pinv;  // -p^-1 mod 2^64
    uint64_t r1;    // R mod p: Montgomery form of 1
    uint64_t r2;    // R^2 mod p: converts a normal word to Montgomery form
    uint64_t ninv;  // n^-1 mod p, normal form (the inverse's scale)
};

// a b R^-1 mod p in [0, p) whenever a b < p R (e.g. a < p and any 64-bit b).
// t = a b; m = t_lo (-p^-1) mod 2^64 makes t + m p divisible by 2^64; the
// quotient is below 2p and may carry out of 64 bits when p > 2^63.
M64_HD uint64_t m64_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t pinv) {
    uint64_t tlo, thi, mlo, mhi;
    gl_mul_wide(a, b, tlo, thi);
    const uint64_t m = tlo * pinv;
    gl_mul_wide(m, p, mlo, mhi);
    (void)mlo;  // tlo + mlo == 0 mod 2^64, with a carry exactly when tlo != 0
    // u = thi + mhi + carry < 2p as one carry chain (the carry enters as its
    // carry-in); one subtraction of p when u carried out or u - p does not
    // borrow
    unsigned c1, c2;
    const uint32_t u0 = __builtin_addc((uint32_t)thi, (uint32_t)mhi, tlo != 0 ? 1u : 0u, &c1);
    const uint32_t u1 = __builtin_addc((uint32_t)(thi >> 32), (uint32_t)(mhi >> 32), c1, &c2);
    const uint64_t u = ((uint64_t)u1 << 32) | u0;
    uint64_t r;
    const bool below = gl_sub_borrow(u, p, r);  // u < p
    return (c2 || !below) ? r : u;
}

// a + b (mod p) for a, b < p: a + b < 2p, so at most one subtraction, due
// when the sum carried out of 64 bits or s - p does not borrow.  Both flags
// come from two-word carry chains.
M64_HD uint64_t m64_add(uint64_t a, uint64_t b, uint64_t p) {
    unsigned c1, c2;
    const uint32_t s0 = __builtin_addc((uint32_t)a, (uint32_t)b, 0u, &c1);
    const uint32_t s1 = __builtin_addc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c1, &c2);
    const uint64_t s = ((uint64_t)s1 << 32) | s0;
    uint64_t r;
    const bool below = gl_sub_borrow(s, p, r);  // s < p
    return (c2 || !below) ? r : s;
}

// a - b (mod p) for a, b < p; the subtraction's own borrow selects + p
M64_HD uint64_t m64_sub(uint64_t a, uint64_t b, uint64_t p) {
    uint64_t d;
    return gl_sub_borrow(a, b, d) ? d + p : d;
}

// host-side constants of a modulus
inline uint64_t m64_neg_inverse(uint64_t p) {
    uint64_t x = p;                                     // p^-1 mod 2^3 for odd p
    for (int i = 0; i < 5; ++i) x *= 2 - p * x;         // Newton: doubles the correct bits
    return (uint64_t)0 - x;
}

}  // namespace stntt
