/*
 * ntt_oracle.c -- CPU restatement of the reference's Stockham NTT.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
 * the product.  Citations are reference file:line.
 */
#include "ntt_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static int is_pow2(uint64_t n) { return n != 0 && (n & (n - 1)) == 0; }

static unsigned ilog2(uint64_t n) {
    unsigned l = 0;
    while ((1ull << l) < n) ++l;
    return l;
}

/* ---------------------------------------------------------------------- */
/* Parameters                                                             */
/* ---------------------------------------------------------------------- */

/* floor(sqrt(n)) by bit search: src/ntt/ntt_util.cpp:72-84 */
uint64_t orc_isqrt(uint64_t n) {
    uint64_t i = 1, result = 0;
    while (i <= n / i) i *= 2;
    for (; i > 0; i /= 2)
        if ((u128)(result + i) * (result + i) <= n) result += i;
    return result;
}

/* trial division over [2, sqrt(n)]: src/ntt/ntt_util.h:65-69.  Like the
   reference, 0 and 1 are reported prime (empty range); callers only ask for
   k*n+1 >= 2. */
int orc_is_prime(uint64_t n) {
    uint64_t end = orc_isqrt(n);
    for (uint64_t x = 2; x <= end; ++x)
        if (n % x == 0) return 0;
    return 1;
}

/* square-and-multiply, positive exponent: src/ntt/ntt_util.h:87-95
   (products widened to 128 bits, so any 64-bit modulus is exact) */
uint64_t orc_powmod(uint64_t base, uint64_t exp, uint64_t mod) {
    uint64_t result = 1 % mod;
    base %= mod;
    while (exp > 0) {
        if (exp & 1) result = (uint64_t)(((u128)result * base) % mod);
        base = (uint64_t)(((u128)base * base) % mod);
        exp >>= 1;
    }
    return result;
}

/* Fermat inverse base^(mod-2): src/ntt/ntt_util.h:74-85 */
uint64_t orc_invmod(uint64_t a, uint64_t mod) { return orc_powmod(a, mod - 2, mod); }

/* smallest prime k*veclen+1 >= min: src/ntt/ntt_util.cpp:16-27 */
uint64_t orc_find_modulus(uint64_t veclen, uint64_t min_mod) {
    uint64_t start = (min_mod - 1 + veclen - 1) / veclen;
    for (uint64_t k = start > 1 ? start : 1;; ++k) {
        if (k > UINT64_MAX / veclen) return 0;
        uint64_t n = k * veclen + 1;
        if (orc_is_prime(n)) return n;
    }
}

/* distinct prime factors by trial division: src/ntt/ntt_util.cpp:51-69 */
static int unique_prime_factors(uint64_t n, uint64_t *out) {
    int cnt = 0;
    uint64_t end = orc_isqrt(n);
    for (uint64_t i = 2; i <= end; i++) {
        if (n % i == 0) {
            n /= i;
            out[cnt++] = i;
            while (n % i == 0) n /= i;
            end = orc_isqrt(n);
        }
    }
    if (n > 1) out[cnt++] = n;
    return cnt;
}

/* src/ntt/ntt_util.cpp:44-49 */
static int is_primitive_root(uint64_t val, uint64_t degree, uint64_t mod) {
    uint64_t f[64];
    int cnt = unique_prime_factors(degree, f);
    if (orc_powmod(val, degree, mod) != 1) return 0;
    for (int i = 0; i < cnt; ++i)
        if (orc_powmod(val, degree / f[i], mod) == 1) return 0;
    return 1;
}

/* first i in [1, mod) that generates Z_mod^*: src/ntt/ntt_util.cpp:36-42 */
uint64_t orc_find_generator(uint64_t totient, uint64_t mod) {
    for (uint64_t i = 1; i < mod; ++i)
        if (is_primitive_root(i, totient, mod)) return i;
    return 0;
}

/* (root, mod) = (g^((p-1)/n), p): src/ntt/ntt_util.cpp:10-14, 29-34 */
int orc_find_params(uint64_t n, uint64_t min_mod, uint64_t *root, uint64_t *mod) {
    if (n == 0) return -2;
    uint64_t p = orc_find_modulus(n, min_mod);
    if (p == 0) return -2;
    uint64_t g = orc_find_generator(p - 1, p);
    if (g == 0) return -2;
    *mod = p;
    *root = orc_powmod(g, (p - 1) / n, p);
    return 0;
}

/* src/ntt/ntt_util.h:35-43 */
size_t orc_bit_reverse(size_t x, unsigned bits) {
    size_t y = 0;
    for (unsigned i = 0; i < bits; ++i) {
        y = (y << 1) | (x & 1);
        x >>= 1;
    }
    return y;
}

/* src/ntt/ntt_util.h:46-57 */
static void bitrev_u32(uint32_t *v, uint64_t n) {
    unsigned lv = ilog2(n);
    for (size_t i = 0; i < n; ++i) {
        size_t j = orc_bit_reverse(i, lv);
        if (j > i) {
            uint32_t t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* 32-bit prime field, exact                                              */
/* ---------------------------------------------------------------------- */

static inline uint32_t mulmod32(uint32_t a, uint32_t b, uint32_t m) {
    return (uint32_t)(((uint64_t)a * b) % m);
}

/* pow_table loop of src/ntt/ntt_cpu.cpp:57-62 (and common.cuh:37-41) */
int orc_twiddles_u32(uint32_t *out, uint64_t n, uint32_t root, uint32_t mod) {
    if (!is_pow2(n)) return -1;
    uint32_t t = 1 % mod;
    for (uint64_t i = 0; i < n / 2; ++i) {
        out[i] = t;
        t = mulmod32(t, root, mod);
    }
    return 0;
}

/* Stockham DIF, natural order in and out: src/ntt/ntt_cpu.cpp:46-85.
   Stage i: J = pow(2, i, mod) (:68; 2^i while 2^i < mod, so for every
   stage below log2 n); for s < n/(2J), w = T[J*s]; for j < J:
     y[2sJ+j]     = x[sJ+j] + x[sJ+j+n/2]
     y[(2s+1)J+j] = w * (x[sJ+j] - x[sJ+j+n/2])
   then swap x and y; an odd number of swaps copies back (:81-84).  y starts
   zeroed (:52) and positions a stage does not write keep their words, which
   matters only for nsteps past log2 n.  Operands are reduced mod p as they
   are read (the reference's modulo() accepts any int). */
int orc_stockham_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod, int nsteps) {
    if (!is_pow2(n)) return -1;
    if (mod < 2) return -2;
    int steps = (int)ilog2(n);
    if (nsteps >= 0) steps = nsteps;
    for (int i = 0; i < steps; i++)
        if (orc_powmod(2, (uint64_t)i, mod) == 0) return -2; /* the reference divides by J */
    uint32_t *y = (uint32_t *)calloc(n, sizeof(uint32_t));
    uint32_t *T = (uint32_t *)malloc((n / 2 + 1) * sizeof(uint32_t));
    if (!y || !T) {
        free(y);
        free(T);
        return -2;
    }
    orc_twiddles_u32(T, n, root, mod);
    uint32_t *x = vec, *orig = vec;
    for (int i = 0; i < steps; i++) {
        uint64_t J = orc_powmod(2, (uint64_t)i, mod);
        for (uint64_t s = 0; s < n / (2 * J); s++) {
            uint32_t w = T[J * s];
            for (uint64_t j = 0; j < J; ++j) {
                uint32_t l = x[s * J + j] % mod, r = x[s * J + j + n / 2] % mod;
                uint32_t sum = (uint32_t)(((uint64_t)l + r) % mod);
                uint32_t dif = (uint32_t)(((uint64_t)l + mod - r) % mod);
                y[2 * s * J + j] = sum;
                y[(2 * s + 1) * J + j] = mulmod32(w, dif, mod);
            }
        }
        uint32_t *t = x;
        x = y;
        y = t;
    }
    if (x != orig) memcpy(orig, x, n * sizeof(uint32_t));
    free(x == orig ? y : x);
    free(T);
    return 0;
}

int orc_stockham_u32_batch(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod, uint64_t batches, int nsteps) {
    for (uint64_t b = 0; b < batches; ++b) {
        int rc = orc_stockham_u32(vec + b * n, n, root, mod, nsteps);
        if (rc) return rc;
    }
    return 0;
}

/* x = n^-1 * DFT_{root^-1}(X).  Equal to the reference's
   inttGs(bitReverse(X)) (src/ntt/ntt_cpu.cpp:130-164); tests check that. */
int orc_inverse_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod) {
    if (!is_pow2(n)) return -1;
    if (n == 1) return 0; /* zero stages: identity, as the forward */
    uint32_t iroot = (uint32_t)orc_invmod(root, mod);
    int rc = orc_stockham_u32(vec, n, iroot, mod, -1);
    if (rc) return rc;
    uint32_t ninv = (uint32_t)orc_invmod(n % mod, mod);
    for (uint64_t i = 0; i < n; ++i) vec[i] = mulmod32(vec[i], ninv, mod);
    return 0;
}

int orc_inverse_u32_batch(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod, uint64_t batches) {
    for (uint64_t b = 0; b < batches; ++b) {
        int rc = orc_inverse_u32(vec + b * n, n, root, mod);
        if (rc) return rc;
    }
    return 0;
}

/* textbook definition, independent of any FFT structure */
int orc_naive_dft_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t root, uint32_t mod) {
    if (n == 0) return -2;
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t wk = orc_powmod(root, k, mod), w = 1 % mod, acc = 0;
        for (uint64_t j = 0; j < n; ++j) {
            acc = (acc + (uint64_t)(in[j] % mod) * w) % mod;
            w = (w * wk) % mod;
        }
        out[k] = (uint32_t)acc;
    }
    return 0;
}

/* Cooley-Tukey, natural in / bit-reversed out: src/ntt/ntt_cpu.cpp:11-42.
   Twiddles root^i (i < n/2) in bit-reversed order (:20-26), k advances once
   per block (:28-41).  The reference leaves vec[l]*T[k] unreduced; here the
   product is exact. */
int orc_ntt_ct_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod) {
    if (!is_pow2(n)) return -1;
    uint32_t *T = (uint32_t *)malloc((n / 2 + 1) * sizeof(uint32_t));
    if (!T) return -2;
    orc_twiddles_u32(T, n, root, mod);
    if (n >= 2) bitrev_u32(T, n / 2);
    for (uint64_t i = 0; i < n; ++i) vec[i] %= mod;
    for (uint64_t size = n; size >= 2; size /= 2) {
        uint64_t half = size / 2, k = 0;
        for (uint64_t i = 0; i < n; i += size) {
            for (uint64_t j = i; j < i + half; ++j) {
                uint64_t l = j + half;
                uint32_t left = vec[j];
                uint32_t right = mulmod32(vec[l], T[k], mod);
                vec[j] = (uint32_t)(((uint64_t)left + right) % mod);
                vec[l] = (uint32_t)(((uint64_t)left + mod - right) % mod);
            }
            k++;
        }
    }
    free(T);
    return 0;
}

/* Gentleman-Sande, bit-reversed in / natural out, scaled by n^-1:
   src/ntt/ntt_cpu.cpp:130-164.  Twiddles root^-i (i < n/2) bit-reversed
   (:137-143); butterfly (l+r, (l-r)*T[k]) (:145-159); scale (:161-163). */
int orc_intt_gs_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod) {
    if (!is_pow2(n)) return -1;
    uint32_t *T = (uint32_t *)malloc((n / 2 + 1) * sizeof(uint32_t));
    if (!T) return -2;
    uint32_t iroot = (uint32_t)orc_invmod(root, mod);
    orc_twiddles_u32(T, n, iroot, mod);
    if (n >= 2) bitrev_u32(T, n / 2);
    for (uint64_t i = 0; i < n; ++i) vec[i] %= mod;
    for (uint64_t size = 2; size <= n; size *= 2) {
        uint64_t half = size / 2, k = 0;
        for (uint64_t i = 0; i < n; i += size) {
            for (uint64_t j = i; j < i + half; ++j) {
                uint64_t l = j + half;
                uint32_t left = vec[j], right = vec[l];
                vec[j] = (uint32_t)(((uint64_t)left + right) % mod);
                vec[l] = mulmod32((uint32_t)(((uint64_t)left + mod - right) % mod), T[k], mod);
            }
            k++;
        }
    }
    uint32_t scaler = (uint32_t)orc_invmod(n % mod, mod);
    for (uint64_t i = 0; i < n; ++i) vec[i] = mulmod32(vec[i], scaler, mod);
    free(T);
    return 0;
}

/* src/ntt/ntt_cpu.cpp:167-174 with elemWiseMul (ntt_util.h:28-33) exact */
int orc_pwc_ntt_u32(uint32_t *a, uint32_t *b, uint64_t n, uint32_t root, uint32_t mod) {
    int rc = orc_ntt_ct_u32(a, n, root, mod);
    if (!rc) rc = orc_ntt_ct_u32(b, n, root, mod);
    if (rc) return rc;
    for (uint64_t i = 0; i < n; ++i) a[i] = mulmod32(a[i], b[i], mod);
    return orc_intt_gs_u32(a, n, root, mod);
}

/* ---------------------------------------------------------------------- */
/* Reference-compatible int32 wrap mode                                   */
/* ---------------------------------------------------------------------- */

/* int*int with two's-complement wrap (what -fwrapv gives the reference) */
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

/* src/ntt/ntt_util.cpp:88-94 */
static inline int32_t ref_modulo(int32_t x, int32_t mod) {
    int32_t r = x % mod;
    return r >= 0 ? r : r + mod;
}

/* pow<int> with positive exponent, int arithmetic: src/ntt/ntt_util.h:87-95 */
static int32_t ref_pow_i32(int32_t base, int32_t exp, int32_t mod) {
    int32_t result = 1;
    base %= mod;
    while (exp > 0) {
        if (exp % 2 == 1) result = wmul(result, base) % mod;
        base = wmul(base, base) % mod;
        exp /= 2;
    }
    return result;
}

/* src/ntt/ntt_cpu.cpp:46-85 verbatim in semantics, including the int
   product wrap at :61 and :74 and J = pow(2, i, mod) at :67. */
int orc_stockham_refcompat(int32_t *vec, uint64_t n, int32_t root, int32_t mod, int nsteps) {
    if (!is_pow2(n)) return -1;
    if (mod < 2) return -2;
    int steps = (int)ilog2(n);
    int32_t *y = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *T = (int32_t *)malloc((n / 2 + 1) * sizeof(int32_t));
    if (!y || !T) {
        free(y);
        free(T);
        return -2;
    }
    int32_t temp = 1;
    for (uint64_t i = 0; i < n / 2; ++i) {
        T[i] = temp;
        temp = wmul(temp, root) % mod;
    }
    if (nsteps >= 0) steps = nsteps;
    int32_t *x = vec, *orig = vec;
    for (int i = 0; i < steps; i++) {
        int64_t J = ref_pow_i32(2, i, mod);
        for (int64_t s = 0; s < (int64_t)n / (2 * J); s++) {
            int32_t w = T[J * s];
            for (int64_t j = 0; j < J; ++j) {
                int32_t left = x[s * J + j];
                int32_t right = x[s * J + j + n / 2];
                y[2 * s * J + j] = ref_modulo((int32_t)((uint32_t)left + (uint32_t)right), mod);
                y[(2 * s + 1) * J + j] =
                    ref_modulo(wmul(w, ref_modulo((int32_t)((uint32_t)left - (uint32_t)right), mod)), mod);
            }
        }
        int32_t *t = x;
        x = y;
        y = t;
    }
    if (x != orig) memcpy(orig, x, n * sizeof(int32_t));
    free(x == orig ? y : x);
    free(T);
    return 0;
}

/* data movement of the Stockham without arithmetic: src/ntt/ntt_cpu.cpp:92-122,
   stage strides J = pow(2, i, mod) (:105) and a zeroed y (:98) */
int orc_stockham_idx_mod(int32_t *vec, uint64_t n, uint32_t mod, int nsteps) {
    if (!is_pow2(n)) return -1;
    if (mod < 2) return -2;
    int steps = (int)ilog2(n);
    if (nsteps >= 0) steps = nsteps;
    for (int i = 0; i < steps; i++)
        if (orc_powmod(2, (uint64_t)i, mod) == 0) return -2; /* the reference divides by J */
    int32_t *y = (int32_t *)calloc(n, sizeof(int32_t));
    if (!y) return -2;
    int32_t *x = vec, *orig = vec;
    for (int i = 0; i < steps; i++) {
        uint64_t J = orc_powmod(2, (uint64_t)i, mod);
        for (uint64_t s = 0; s < n / (2 * J); s++)
            for (uint64_t j = 0; j < J; ++j) {
                y[2 * s * J + j] = x[s * J + j];
                y[(2 * s + 1) * J + j] = x[s * J + j + n / 2];
            }
        int32_t *t = x;
        x = y;
        y = t;
    }
    if (x != orig) memcpy(orig, x, n * sizeof(int32_t));
    free(x == orig ? y : x);
    return 0;
}

/* with the modulus the reference's own fixtures used (2^31 - 1 > n: J = 2^i) */
int orc_stockham_idx(int32_t *vec, uint64_t n, int nsteps) { return orc_stockham_idx_mod(vec, n, 2147483647u, nsteps); }

/* ---------------------------------------------------------------------- */
/* Goldilocks, exact.  The reference declares the field                   */
/* (src/cuda/ntt/implementations/goldilocks/arithmetic.cuh:16-35) but     */
/* leaves mul/butterfly empty; the arithmetic here is the textbook field  */
/* Z_P with 128-bit products.                                             */
/* ---------------------------------------------------------------------- */

uint64_t orc_gl_mul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % ORC_GL_P); }
uint64_t orc_gl_add(uint64_t a, uint64_t b) { return (uint64_t)(((u128)(a % ORC_GL_P) + (b % ORC_GL_P)) % ORC_GL_P); }
uint64_t orc_gl_sub(uint64_t a, uint64_t b) {
    return (uint64_t)(((u128)(a % ORC_GL_P) + ORC_GL_P - (b % ORC_GL_P)) % ORC_GL_P);
}
uint64_t orc_gl_pow(uint64_t a, uint64_t e) { return orc_powmod(a, e, ORC_GL_P); }

uint64_t orc_gl_root(unsigned log2n) {
    if (log2n > 32) return 0;
    return orc_gl_pow(7, (ORC_GL_P - 1) >> log2n);
}

int orc_twiddles_gl(uint64_t *out, unsigned log2n, int inverse) {
    if (log2n > 32) return -2;
    uint64_t n = 1ull << log2n, w = orc_gl_root(log2n), t = 1;
    if (inverse) w = orc_invmod(w, ORC_GL_P);
    for (uint64_t i = 0; i < n / 2; ++i) {
        out[i] = t;
        t = orc_gl_mul(t, w);
    }
    return 0;
}

/* same stage loop as orc_stockham_u32 (src/ntt/ntt_cpu.cpp:46-85) */
static int stockham_gl_root(uint64_t *vec, uint64_t n, uint64_t root, int nsteps) {
    int steps = (int)ilog2(n);
    if (nsteps >= 0) steps = nsteps;
    uint64_t *y = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *T = (uint64_t *)malloc((n / 2 + 1) * sizeof(uint64_t));
    if (!y || !T) {
        free(y);
        free(T);
        return -2;
    }
    uint64_t t = 1;
    for (uint64_t i = 0; i < n / 2; ++i) {
        T[i] = t;
        t = orc_gl_mul(t, root);
    }
    uint64_t *x = vec, *orig = vec;
    if (steps > 0)
        for (uint64_t i = 0; i < n; ++i) x[i] %= ORC_GL_P;
    for (int i = 0; i < steps; i++) {
        uint64_t J = 1ull << i;
        for (uint64_t s = 0; s < n / (2 * J); s++) {
            uint64_t w = T[J * s];
            for (uint64_t j = 0; j < J; ++j) {
                uint64_t l = x[s * J + j], r = x[s * J + j + n / 2];
                y[2 * s * J + j] = orc_gl_add(l, r);
                y[(2 * s + 1) * J + j] = orc_gl_mul(w, orc_gl_sub(l, r));
            }
        }
        uint64_t *tp = x;
        x = y;
        y = tp;
    }
    if (x != orig) memcpy(orig, x, n * sizeof(uint64_t));
    free(x == orig ? y : x);
    free(T);
    return 0;
}

int orc_stockham_gl(uint64_t *vec, unsigned log2n, int nsteps) {
    if (log2n > 32) return -2;
    return stockham_gl_root(vec, 1ull << log2n, orc_gl_root(log2n), nsteps);
}

int orc_stockham_gl_batch(uint64_t *vec, unsigned log2n, uint64_t batches, int nsteps) {
    for (uint64_t b = 0; b < batches; ++b) {
        int rc = orc_stockham_gl(vec + (b << log2n), log2n, nsteps);
        if (rc) return rc;
    }
    return 0;
}

int orc_inverse_gl(uint64_t *vec, unsigned log2n) {
    if (log2n > 32) return -2;
    if (log2n == 0) return 0; /* zero stages: identity, as the forward */
    uint64_t n = 1ull << log2n;
    int rc = stockham_gl_root(vec, n, orc_invmod(orc_gl_root(log2n), ORC_GL_P), -1);
    if (rc) return rc;
    uint64_t ninv = orc_invmod(n % ORC_GL_P, ORC_GL_P);
    for (uint64_t i = 0; i < n; ++i) vec[i] = orc_gl_mul(vec[i], ninv);
    return 0;
}

int orc_inverse_gl_batch(uint64_t *vec, unsigned log2n, uint64_t batches) {
    for (uint64_t b = 0; b < batches; ++b) {
        int rc = orc_inverse_gl(vec + (b << log2n), log2n);
        if (rc) return rc;
    }
    return 0;
}

static void bitrev_u64(uint64_t *v, uint64_t n) {
    unsigned lv = ilog2(n);
    for (size_t i = 0; i < n; ++i) {
        size_t j = orc_bit_reverse(i, lv);
        if (j > i) {
            uint64_t t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
}

/* Goldilocks twin of orc_ntt_ct_u32: the loop of src/ntt/ntt_cpu.cpp:11-42
   (bit-reversed table of w^i, i < n/2; right = v[l] T[k]; k per block) */
int orc_ntt_ct_gl(uint64_t *vec, unsigned log2n) {
    if (log2n > 32) return -2;
    uint64_t n = 1ull << log2n;
    uint64_t *T = (uint64_t *)malloc((n / 2 + 1) * sizeof(uint64_t));
    if (!T) return -2;
    orc_twiddles_gl(T, log2n, 0);
    if (n >= 2) bitrev_u64(T, n / 2);
    for (uint64_t i = 0; i < n; ++i) vec[i] %= ORC_GL_P;
    for (uint64_t size = n; size >= 2; size /= 2) {
        uint64_t half = size / 2, k = 0;
        for (uint64_t i = 0; i < n; i += size) {
            for (uint64_t j = i; j < i + half; ++j) {
                uint64_t l = j + half;
                uint64_t left = vec[j], right = orc_gl_mul(vec[l], T[k]);
                vec[j] = orc_gl_add(left, right);
                vec[l] = orc_gl_sub(left, right);
            }
            k++;
        }
    }
    free(T);
    return 0;
}

/* Goldilocks twin of orc_intt_gs_u32: src/ntt/ntt_cpu.cpp:130-164 */
int orc_intt_gs_gl(uint64_t *vec, unsigned log2n) {
    if (log2n > 32) return -2;
    uint64_t n = 1ull << log2n;
    uint64_t *T = (uint64_t *)malloc((n / 2 + 1) * sizeof(uint64_t));
    if (!T) return -2;
    orc_twiddles_gl(T, log2n, 1);
    if (n >= 2) bitrev_u64(T, n / 2);
    for (uint64_t i = 0; i < n; ++i) vec[i] %= ORC_GL_P;
    for (uint64_t size = 2; size <= n; size *= 2) {
        uint64_t half = size / 2, k = 0;
        for (uint64_t i = 0; i < n; i += size) {
            for (uint64_t j = i; j < i + half; ++j) {
                uint64_t l = j + half;
                uint64_t left = vec[j], right = vec[l];
                vec[j] = orc_gl_add(left, right);
                vec[l] = orc_gl_mul(orc_gl_sub(left, right), T[k]);
            }
            k++;
        }
    }
    uint64_t scaler = orc_invmod(n % ORC_GL_P, ORC_GL_P);
    for (uint64_t i = 0; i < n; ++i) vec[i] = orc_gl_mul(vec[i], scaler);
    free(T);
    return 0;
}

/* Goldilocks twin of orc_pwc_ntt_u32 (src/ntt/ntt_cpu.cpp:167-174) */
int orc_pwc_ntt_gl(uint64_t *a, uint64_t *b, unsigned log2n) {
    int rc = orc_ntt_ct_gl(a, log2n);
    if (!rc) rc = orc_ntt_ct_gl(b, log2n);
    if (rc) return rc;
    for (uint64_t i = 0; i < (1ull << log2n); ++i) a[i] = orc_gl_mul(a[i], b[i]);
    return orc_intt_gs_gl(a, log2n);
}

int orc_naive_dft_gl(const uint64_t *in, uint64_t *out, unsigned log2n, int inverse) {
    if (log2n > 32) return -2;
    uint64_t n = 1ull << log2n, root = orc_gl_root(log2n);
    if (inverse) root = orc_invmod(root, ORC_GL_P);
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t wk = orc_gl_pow(root, k), w = 1, acc = 0;
        for (uint64_t j = 0; j < n; ++j) {
            acc = orc_gl_add(acc, orc_gl_mul(in[j] % ORC_GL_P, w));
            w = orc_gl_mul(w, wk);
        }
        out[k] = acc;
    }
    if (inverse) {
        uint64_t ninv = orc_invmod(n % ORC_GL_P, ORC_GL_P);
        for (uint64_t k = 0; k < n; ++k) out[k] = orc_gl_mul(out[k], ninv);
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Single outputs of the textbook DFT, O(n) each: X[k] = sum_j x[j] w^{jk} */
/* (times n^-1 for the inverse).  Used to check sampled outputs of large   */
/* transforms without running a full O(n log n) oracle.                   */
/* ---------------------------------------------------------------------- */

/* x mod P for a 128-bit x, special form: 2^64 == 2^32-1, 2^96 == -1 */
static inline uint64_t gl_reduce_fast(u128 x) {
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t h0 = hi & 0xFFFFFFFFull, h1 = hi >> 32;
    u128 t = (u128)lo + (u128)h0 * 0xFFFFFFFFull + ORC_GL_P - h1;
    return (uint64_t)(t % ORC_GL_P);
}

int orc_dft_points_gl(const uint64_t *x, unsigned log2n, int inverse, const uint64_t *ks, uint64_t count,
                      uint64_t *out) {
    if (log2n > 32) return -2;
    uint64_t n = 1ull << log2n, root = orc_gl_root(log2n);
    if (inverse) root = orc_invmod(root, ORC_GL_P);
    uint64_t ninv = inverse ? orc_invmod(n % ORC_GL_P, ORC_GL_P) : 1;
    for (uint64_t c = 0; c < count; ++c) {
        uint64_t wk = orc_gl_pow(root, ks[c] % n), w = 1;
        u128 acc = 0;
        for (uint64_t j = 0; j < n; ++j) {
            acc += gl_reduce_fast((u128)(x[j] % ORC_GL_P) * w);
            if ((j & 1023) == 1023) acc %= ORC_GL_P;
            w = gl_reduce_fast((u128)w * wk);
        }
        out[c] = orc_gl_mul((uint64_t)(acc % ORC_GL_P), ninv);
    }
    return 0;
}

int orc_dft_points_u32(const uint32_t *x, uint64_t n, uint32_t root, uint32_t mod, int inverse, const uint64_t *ks,
                       uint64_t count, uint32_t *out) {
    if (inverse) root = (uint32_t)orc_invmod(root, mod);
    uint64_t ninv = inverse ? orc_invmod(n % mod, mod) : 1;
    for (uint64_t c = 0; c < count; ++c) {
        uint64_t wk = orc_powmod(root, ks[c] % n, mod), w = 1 % mod, acc = 0;
        for (uint64_t j = 0; j < n; ++j) {
            acc = (acc + (uint64_t)(x[j] % mod) * w) % mod;
            w = w * wk % mod;
        }
        out[c] = (uint32_t)(acc * ninv % mod);
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Any odd modulus below 2^64 (the 64-bit field): the same loops as the   */
/* 32-bit functions above with 128-bit products.                          */
/* ---------------------------------------------------------------------- */
static inline uint64_t mm64_mul(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a * b) % m); }
static inline uint64_t mm64_add(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a + b) % m); }
static inline uint64_t mm64_sub(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a + m - b) % m); }

int orc_twiddles_u64(uint64_t *out, uint64_t n, uint64_t root, uint64_t mod) {
    uint64_t t = 1 % mod;
    for (uint64_t i = 0; i < n / 2; ++i) {
        out[i] = t;
        t = mm64_mul(t, root, mod);
    }
    return 0;
}

/* src/ntt/ntt_cpu.cpp:46-85 over Z_mod */
int orc_stockham_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod, int nsteps) {
    if (!is_pow2(n)) return -1;
    int steps = (int)ilog2(n);
    if (nsteps >= 0) steps = nsteps;
    uint64_t *y = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *T = (uint64_t *)malloc((n / 2 + 1) * sizeof(uint64_t));
    if (!y || !T) {
        free(y);
        free(T);
        return -2;
    }
    orc_twiddles_u64(T, n, root % mod, mod);
    uint64_t *x = vec, *orig = vec;
    if (steps > 0)
        for (uint64_t i = 0; i < n; ++i) x[i] %= mod;
    for (int i = 0; i < steps; i++) {
        uint64_t J = 1ull << i;
        for (uint64_t s = 0; s < n / (2 * J); s++) {
            uint64_t w = T[J * s];
            for (uint64_t j = 0; j < J; ++j) {
                uint64_t l = x[s * J + j], r = x[s * J + j + n / 2];
                y[2 * s * J + j] = mm64_add(l, r, mod);
                y[(2 * s + 1) * J + j] = mm64_mul(w, mm64_sub(l, r, mod), mod);
            }
        }
        uint64_t *tp = x;
        x = y;
        y = tp;
    }
    if (x != orig) memcpy(orig, x, n * sizeof(uint64_t));
    free(x == orig ? y : x);
    free(T);
    return 0;
}

int orc_stockham_u64_batch(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod, uint64_t batches, int nsteps) {
    for (uint64_t b = 0; b < batches; ++b) {
        int rc = orc_stockham_u64(vec + b * n, n, root, mod, nsteps);
        if (rc) return rc;
    }
    return 0;
}

/* natural-order inverse: the forward loop with root^-1, then n^-1 */
int orc_inverse_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod) {
    if (!is_pow2(n)) return -1;
    if (n == 1) {
        vec[0] %= mod;
        return 0;
    }
    int rc = orc_stockham_u64(vec, n, orc_invmod(root % mod, mod), mod, -1);
    if (rc) return rc;
    uint64_t ninv = orc_invmod(n % mod, mod);
    for (uint64_t i = 0; i < n; ++i) vec[i] = mm64_mul(vec[i], ninv, mod);
    return 0;
}

int orc_inverse_u64_batch(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod, uint64_t batches) {
    for (uint64_t b = 0; b < batches; ++b) {
        int rc = orc_inverse_u64(vec + b * n, n, root, mod);
        if (rc) return rc;
    }
    return 0;
}

int orc_naive_dft_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t root, uint64_t mod) {
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t wk = orc_powmod(root % mod, k, mod), w = 1 % mod, acc = 0;
        for (uint64_t j = 0; j < n; ++j) {
            acc = mm64_add(acc, mm64_mul(in[j] % mod, w, mod), mod);
            w = mm64_mul(w, wk, mod);
        }
        out[k] = acc;
    }
    return 0;
}

/* src/ntt/ntt_cpu.cpp:11-42 over Z_mod */
int orc_ntt_ct_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod) {
    if (!is_pow2(n)) return -1;
    uint64_t *T = (uint64_t *)malloc((n / 2 + 1) * sizeof(uint64_t));
    if (!T) return -2;
    orc_twiddles_u64(T, n, root % mod, mod);
    if (n >= 2) bitrev_u64(T, n / 2);
    for (uint64_t i = 0; i < n; ++i) vec[i] %= mod;
    for (uint64_t size = n; size >= 2; size /= 2) {
        uint64_t half = size / 2, k = 0;
        for (uint64_t i = 0; i < n; i += size) {
            for (uint64_t j = i; j < i + half; ++j) {
                uint64_t l = j + half;
                uint64_t left = vec[j], right = mm64_mul(vec[l], T[k], mod);
                vec[j] = mm64_add(left, right, mod);
                vec[l] = mm64_sub(left, right, mod);
            }
            k++;
        }
    }
    free(T);
    return 0;
}

/* src/ntt/ntt_cpu.cpp:130-164 over Z_mod */
int orc_intt_gs_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod) {
    if (!is_pow2(n)) return -1;
    uint64_t *T = (uint64_t *)malloc((n / 2 + 1) * sizeof(uint64_t));
    if (!T) return -2;
    orc_twiddles_u64(T, n, orc_invmod(root % mod, mod), mod);
    if (n >= 2) bitrev_u64(T, n / 2);
    for (uint64_t i = 0; i < n; ++i) vec[i] %= mod;
    for (uint64_t size = 2; size <= n; size *= 2) {
        uint64_t half = size / 2, k = 0;
        for (uint64_t i = 0; i < n; i += size) {
            for (uint64_t j = i; j < i + half; ++j) {
                uint64_t l = j + half;
                uint64_t left = vec[j], right = vec[l];
                vec[j] = mm64_add(left, right, mod);
                vec[l] = mm64_mul(mm64_sub(left, right, mod), T[k], mod);
            }
            k++;
        }
    }
    uint64_t scaler = orc_invmod(n % mod, mod);
    for (uint64_t i = 0; i < n; ++i) vec[i] = mm64_mul(vec[i], scaler, mod);
    free(T);
    return 0;
}

/* src/ntt/ntt_cpu.cpp:167-174 over Z_mod */
int orc_pwc_ntt_u64(uint64_t *a, uint64_t *b, uint64_t n, uint64_t root, uint64_t mod) {
    int rc = orc_ntt_ct_u64(a, n, root, mod);
    if (!rc) rc = orc_ntt_ct_u64(b, n, root, mod);
    if (rc) return rc;
    for (uint64_t i = 0; i < n; ++i) a[i] = mm64_mul(a[i], b[i], mod);
    return orc_intt_gs_u64(a, n, root, mod);
}
