/*
 * ntt_oracle.h -- CPU restatement of the reference's Stockham NTT.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path (stntt_amd/csrc): only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it and
 * has no CPU fallback.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference root).  All arithmetic is exact (64/128-bit products)
 * unless the name says "refcompat", in which case the reference's int32
 * wrap-around is reproduced on purpose (src/ntt/ntt_cpu.cpp:61,74 compute
 * int*int products that overflow for p >= 46341).
 *
 * Return codes: 0 = ok, -1 = size not a power of two (the reference throws
 * std::invalid_argument, src/ntt/ntt_cpu.cpp:48-50), -2 = bad argument.
 */
#ifndef STNTT_ORACLE_H
#define STNTT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_GL_P 0xFFFFFFFF00000001ull /* goldilocks/arithmetic.cuh:16 */

/* ---- parameters: src/ntt/ntt_util.cpp:10-84, ntt_util.h:65-96 ---------- */
uint64_t orc_isqrt(uint64_t n);
int orc_is_prime(uint64_t n);
uint64_t orc_powmod(uint64_t base, uint64_t exp, uint64_t mod);
uint64_t orc_invmod(uint64_t a, uint64_t mod);
uint64_t orc_find_modulus(uint64_t veclen, uint64_t min_mod);
uint64_t orc_find_generator(uint64_t totient, uint64_t mod);
int orc_find_params(uint64_t n, uint64_t min_mod, uint64_t *root, uint64_t *mod);
size_t orc_bit_reverse(size_t x, unsigned bits);

/* ---- 32-bit prime field, exact ---------------------------------------- */
/* twiddle table T[i] = root^i, i < n/2 (common.cuh:37-41, ntt_cpu.cpp:57-62) */
int orc_twiddles_u32(uint32_t *out, uint64_t n, uint32_t root, uint32_t mod);
/* Stockham DIF natural->natural; nsteps < 0 = all (ntt_cpu.cpp:46-85) */
int orc_stockham_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod, int nsteps);
/* batched form (ntt_cpu.cpp:87-90) */
int orc_stockham_u32_batch(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod, uint64_t batches, int nsteps);
/* inverse: n^-1 * DFT_{root^-1}, natural->natural */
int orc_inverse_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod);
int orc_inverse_u32_batch(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod, uint64_t batches);
/* O(n^2) textbook DFT X[k] = sum_j x[j] w^{jk} */
int orc_naive_dft_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t root, uint32_t mod);
/* Cooley-Tukey natural->bit-reversed (ntt_cpu.cpp:11-42), exact */
int orc_ntt_ct_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod);
/* Gentleman-Sande bit-reversed->natural inverse incl. n^-1 (ntt_cpu.cpp:130-164), exact */
int orc_intt_gs_u32(uint32_t *vec, uint64_t n, uint32_t root, uint32_t mod);
/* positive-wrapped (cyclic) convolution through CT / pointwise / GS
   (ntt_cpu.cpp:167-174); result in a */
int orc_pwc_ntt_u32(uint32_t *a, uint32_t *b, uint64_t n, uint32_t root, uint32_t mod);

/* ---- reference-compatible int32 wrap mode ------------------------------ */
/* Stockham exactly as ntt_cpu.cpp:46-85 with int32 overflow made explicit */
int orc_stockham_refcompat(int32_t *vec, uint64_t n, int32_t root, int32_t mod, int nsteps);
/* index-only Stockham (ntt_cpu.cpp:92-122), strides pow(2, i, mod) */
int orc_stockham_idx_mod(int32_t *vec, uint64_t n, uint32_t mod, int nsteps);
/* the same with mod = 2^31 - 1 (strides 2^i) */
int orc_stockham_idx(int32_t *vec, uint64_t n, int nsteps);

/* ---- Goldilocks p = 2^64 - 2^32 + 1, exact ----------------------------- */
uint64_t orc_gl_mul(uint64_t a, uint64_t b);
uint64_t orc_gl_add(uint64_t a, uint64_t b);
uint64_t orc_gl_sub(uint64_t a, uint64_t b);
uint64_t orc_gl_pow(uint64_t a, uint64_t e);
/* primitive n-th root 7^((P-1)/n); 7 = smallest generator (ntt_util.cpp:36-42) */
uint64_t orc_gl_root(unsigned log2n);
int orc_twiddles_gl(uint64_t *out, unsigned log2n, int inverse);
int orc_stockham_gl(uint64_t *vec, unsigned log2n, int nsteps);
int orc_stockham_gl_batch(uint64_t *vec, unsigned log2n, uint64_t batches, int nsteps);
int orc_inverse_gl(uint64_t *vec, unsigned log2n);
int orc_inverse_gl_batch(uint64_t *vec, unsigned log2n, uint64_t batches);
int orc_naive_dft_gl(const uint64_t *in, uint64_t *out, unsigned log2n, int inverse);
/* Goldilocks twins of the CT / GS / convolution loops above (same
   ntt_cpu.cpp lines, Goldilocks arithmetic) */
int orc_ntt_ct_gl(uint64_t *vec, unsigned log2n);
int orc_intt_gs_gl(uint64_t *vec, unsigned log2n);
int orc_pwc_ntt_gl(uint64_t *a, uint64_t *b, unsigned log2n);

/* ---- any odd modulus below 2^64, exact (128-bit products) ------------- */
int orc_twiddles_u64(uint64_t *out, uint64_t n, uint64_t root, uint64_t mod);
int orc_stockham_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod, int nsteps);
int orc_stockham_u64_batch(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod, uint64_t batches, int nsteps);
int orc_inverse_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod);
int orc_inverse_u64_batch(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod, uint64_t batches);
int orc_naive_dft_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t root, uint64_t mod);
int orc_ntt_ct_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod);
int orc_intt_gs_u64(uint64_t *vec, uint64_t n, uint64_t root, uint64_t mod);
int orc_pwc_ntt_u64(uint64_t *a, uint64_t *b, uint64_t n, uint64_t root, uint64_t mod);

/* ---- sampled outputs of the textbook DFT, O(n) per output ------------- */
int orc_dft_points_gl(const uint64_t *x, unsigned log2n, int inverse, const uint64_t *ks, uint64_t count,
                      uint64_t *out);
int orc_dft_points_u32(const uint32_t *x, uint64_t n, uint32_t root, uint32_t mod, int inverse, const uint64_t *ks,
                       uint64_t count, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
