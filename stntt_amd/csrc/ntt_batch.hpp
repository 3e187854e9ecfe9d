// ntt_batch.hpp -- one persistent launch for thousands of independent
// transforms of mixed size, field and direction.
//
// The reference runs one transform (or one same-size batch) per autoNtt call
// (src/cuda/ntt/ntt_kernel.cuh:152-201), each with its own allocation,
// twiddle upload and kernel chain.  Here a descriptor array lists every
// transform; a fixed grid of resident workgroups pulls descriptors from an
// atomic counter (largest first, sorted on the host) until the queue is
// empty, and each workgroup runs its transform to completion:
//   * n * w <= 32 KiB: the whole transform lives in LDS; radix-16 (and one
//     smaller) Stockham sub-steps run in registers between LDS exchanges;
//   * larger n: the same sub-steps stream through global memory, ping-
//     ponging between the destination and a scratch slice private to the
//     workgroup, with a workgroup barrier and workgroup-scope fences between
//     sub-steps.
// No workgroup ever waits for another, so the kernel needs no inter-
// workgroup synchronisation and every wave reaches the exit when the counter
// passes the last descriptor.  Sub-step algebra: see ntt_pass.hpp.
#pragma once

#include "ntt_pass.hpp"

namespace stntt {

constexpr int BATCH_THREADS = 256;
constexpr int BATCH_LDS_BYTES = 32768;

// One transform.  Field parameters and twiddle factors come from the
// context's resident plan for (field, n, root, mod, direction).
struct BatchTask {
    const void *src;
    void *dst;
    const void *hi;  // hi[e] = w^((n / Rmax) e)
    const void *lo;  // lo[e] = w^e, e < n / Rmax
    uint32_t field;  // 0 = 32-bit prime field, 1 = Goldilocks
    uint32_t dir;
    uint32_t log_n;
    uint32_t log_rmax;
    uint32_t flags;  // PASS_* bits for the first / last sub-step
    uint32_t nsub;   // number of sub-steps
    uint32_t radix_logs;  // 4 bits per sub-step, first sub-step in the low bits
    uint32_t pad;
    FieldU32::Rt rt32;
    FieldGL::Rt rtgl;
};

// w^e (e < n) from the two factor tables
template <class FP>
__device__ __forceinline__ typename FP::T batch_twiddle_mul(typename FP::T x, uint32_t e, const BatchTask &t,
                                                            const typename FP::Rt &rt) {
    const typename FP::Tw *hi = (const typename FP::Tw *)t.hi;
    const typename FP::Tw *lo = (const typename FP::Tw *)t.lo;
    const uint32_t log_b = t.log_n - t.log_rmax;
    x = FP::mul(x, at_off(hi, e >> log_b), rt);
    if (log_b) x = FP::mul(x, at_off(lo, e & ((1u << log_b) - 1)), rt);
    return x;
}

// Output k of one radix-r group: the sub-step twiddle w_n^(s J' k) (none on
// the last sub-step, where s = 0), then the finishing touch of the last
// pass (inverse scaling).
template <class FP>
__device__ __forceinline__ typename FP::T batch_out(typename FP::T x, uint32_t s, uint32_t logj, int k,
                                                   const BatchTask &t, const typename FP::Rt &rt, uint32_t flags,
                                                   bool last) {
    if (!last && k != 0) x = batch_twiddle_mul<FP>(x, (s << logj) * k, t, rt);
    if (last && (flags & PASS_LAST)) x = FP::finish(x, rt, flags);
    return x;
}

// One radix-2^LR Stockham sub-step over a whole LDS-resident transform of
// n = 2^L points, in place: every thread reads all its groups, the
// workgroup synchronises, then writes.  Inner J' = 2^logj.
template <class FP, int LR, int DIR, int GMAX>
__device__ __forceinline__ void batch_substep_lds(shared_t<typename FP::T> *z, uint32_t L, uint32_t logj,
                                                  const BatchTask &t,
                                                  const typename FP::Rt &rt, uint32_t flags, bool last) {
    using T = typename FP::T;
    constexpr int r = 1 << LR;
    const uint32_t ngroups = 1u << (L - LR);  // also the input stride n / r
    const auto roots = FP::template load_roots<LR, DIR>((const typename FP::Tw *)t.hi, t.log_rmax);
    T v[GMAX][r];
#pragma unroll
    for (int gi = 0; gi < GMAX; ++gi) {
        const uint32_t g = threadIdx.x + gi * BATCH_THREADS;
        if (g < ngroups) {
#pragma unroll
            for (int m = 0; m < r; ++m) v[gi][m] = z[g + m * ngroups];
        }
    }
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < GMAX; ++gi) {
        const uint32_t g = threadIdx.x + gi * BATCH_THREADS;
        if (g < ngroups) {
            dft_regs<FP, LR, DIR>(v[gi], roots, rt);
            const uint32_t s = g >> logj, j = g & ((1u << logj) - 1);
#pragma unroll
            for (int k = 0; k < r; ++k)
                z[((r * s + k) << logj) + j] = batch_out<FP>(v[gi][bitrev_c(k, LR)], s, logj, k, t, rt, flags, last);
        }
    }
}

// The same sub-step streaming from `in` to a distinct `out` in global
// memory, one group at a time.
template <class FP, int LR, int DIR>
__device__ __forceinline__ void batch_substep_global(const global_t<typename FP::T> *in,
                                                     global_t<typename FP::T> *out, uint32_t L,
                                                     uint32_t logj, const BatchTask &t, const typename FP::Rt &rt,
                                                     uint32_t flags, bool last) {
    using T = typename FP::T;
    constexpr int r = 1 << LR;
    const uint32_t ngroups = 1u << (L - LR);
    const auto roots = FP::template load_roots<LR, DIR>((const typename FP::Tw *)t.hi, t.log_rmax);
    for (uint32_t g = threadIdx.x; g < ngroups; g += BATCH_THREADS) {
        T v[r];
#pragma unroll
        for (int m = 0; m < r; ++m) v[m] = FP::load_in(in[g + m * ngroups], rt, flags);
        dft_regs<FP, LR, DIR>(v, roots, rt);
        const uint32_t s = g >> logj, j = g & ((1u << logj) - 1);
#pragma unroll
        for (int k = 0; k < r; ++k)
            out[((r * s + k) << logj) + j] = batch_out<FP>(v[bitrev_c(k, LR)], s, logj, k, t, rt, flags, last);
    }
}

template <class FP, int DIR>
__device__ __forceinline__ void batch_run_task(const BatchTask &t, typename FP::T *scratch, unsigned char *smem) {
    using T = typename FP::T;
    const typename FP::Rt rt = [&] {
        if constexpr (std::is_same_v<FP, FieldGL>) return t.rtgl;
        else return t.rt32;
    }();
    const uint32_t L = t.log_n;
    const uint32_t n = 1u << L;
    // descriptor pointers are device memory: global_t keeps the accesses
    // global (not flat) so LDS waits do not also wait for HBM stores
    const global_t<T> *src = (const global_t<T> *)t.src;
    global_t<T> *dst = (global_t<T> *)t.dst;
    global_t<T> *gscratch = (global_t<T> *)scratch;
    if (L == 0) {  // zero stages: identity
        if (threadIdx.x == 0) dst[0] = src[0];
        return;
    }
    if (n * sizeof(T) <= (uint32_t)BATCH_LDS_BYTES) {
        // ---- whole transform in LDS ----
        shared_t<T> *z = (shared_t<T> *)smem;  // LDS accesses, not flat
        for (uint32_t i = threadIdx.x; i < n; i += BATCH_THREADS) z[i] = FP::load_in(src[i], rt, t.flags);
        __syncthreads();
        uint32_t logj = 0;
        for (uint32_t i = 0; i < t.nsub; ++i) {
            const int lr = (t.radix_logs >> (4 * i)) & 15;
            const bool last = i + 1 == t.nsub;
            // at most 2^13 points (32 KiB of 32-bit words): <= 2^(13-lr) groups
            const uint32_t fl = t.flags & ~PASS_FIRST;
            switch (lr) {
                case 4: batch_substep_lds<FP, 4, DIR, 2>(z, L, logj, t, rt, fl, last); break;
                case 3: batch_substep_lds<FP, 3, DIR, 4>(z, L, logj, t, rt, fl, last); break;
                case 2: batch_substep_lds<FP, 2, DIR, 8>(z, L, logj, t, rt, fl, last); break;
                default: batch_substep_lds<FP, 1, DIR, 16>(z, L, logj, t, rt, fl, last); break;
            }
            __syncthreads();
            logj += lr;
        }
        for (uint32_t i = threadIdx.x; i < n; i += BATCH_THREADS) dst[i] = z[i];
        __syncthreads();
    } else {
        // ---- sub-steps through global memory: src -> ... -> dst ----
        // buffers alternate so that the last sub-step writes dst; the first
        // reads src (which may equal dst only when the count is even)
        uint32_t logj = 0;
        const global_t<T> *in = src;
        for (uint32_t i = 0; i < t.nsub; ++i) {
            const int lr = (t.radix_logs >> (4 * i)) & 15;
            const bool last = i + 1 == t.nsub;
            global_t<T> *out = ((t.nsub - 1 - i) & 1) ? gscratch : dst;
            const uint32_t fl = i == 0 ? t.flags : (t.flags & ~PASS_FIRST);
            switch (lr) {
                case 4: batch_substep_global<FP, 4, DIR>(in, out, L, logj, t, rt, fl, last); break;
                case 3: batch_substep_global<FP, 3, DIR>(in, out, L, logj, t, rt, fl, last); break;
                case 2: batch_substep_global<FP, 2, DIR>(in, out, L, logj, t, rt, fl, last); break;
                default: batch_substep_global<FP, 1, DIR>(in, out, L, logj, t, rt, fl, last); break;
            }
            // Make this workgroup's global writes visible to its own next
            // reads.  Workgroup scope suffices: no other workgroup touches
            // this transform's buffers during the launch, and the waves of a
            // workgroup share one CU's caches.  (Agent scope would write back
            // and invalidate the whole L2 of the XCD at every sub-step.)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            in = out;
            logj += lr;
        }
    }
}

}  // namespace stntt
