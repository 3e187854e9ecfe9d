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
    (void)mlo;                         // tlo + mlo == 0 mod 2^64
    const uint64_t c = tlo != 0;       // carry of tlo + mlo
    uint64_t u = thi + mhi;
    bool ov = u < thi;
    const uint64_t u2 = u + c;
    ov = ov || u2 < u;
    const uint64_t r = u2 - p;
    return (ov || u2 >= p) ? r : u2;
}

M64_HD uint64_t m64_add(uint64_t a, uint64_t b, uint64_t p) {
    const uint64_t s = a + b;  // a + b < 2p: at most one subtraction
    const uint64_t r = s - p;
    return (s < a || s >= p) ? r : s;
}

M64_HD uint64_t m64_sub(uint64_t a, uint64_t b, uint64_t p) {
    const uint64_t d = a - b;
    return a < b ? d + p : d;
}

// host-side constants of a modulus
inline uint64_t m64_neg_inverse(uint64_t p) {
    uint64_t x = p;                                     // p^-1 mod 2^3 for odd p
    for (int i = 0; i < 5; ++i) x *= 2 - p * x;         // Newton: doubles the correct bits
    return (uint64_t)0 - x;
}

}  // namespace stntt
