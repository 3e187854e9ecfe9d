// launch.hpp -- host-side launchers of the transform kernels.
//
// The pass kernels of each (field, direction) and the persistent batch
// kernel are instantiated in translation units of their own
// (launch_pass.hip, compiled once per field and direction, and
// batch_kernel.hip), so the library's device code compiles in parallel.
// Each launcher enqueues on stream `s` and returns without synchronising;
// errors surface through hipGetLastError() as for a direct launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdlib.h>

#include <type_traits>

#include "ntt_batch.hpp"
#include "ntt_big.hpp"
#include "ntt_pass.hpp"

namespace stntt {

// One pass of a transform schedule (see plan.hpp):
//   single: the whole transform in one LDS pass (R = n), `narrow` selecting
//           the one-transform-per-workgroup variant;
//   otherwise a tile pass of radix 2^logr; `first` (J = 1) / last pass are
//           told apart by `first` and PassArgs::flags; `perm` stages a
//           bit-reversed input (first pass) or output (last pass).
template <class FP, int DIR>
void launch_pass_dir(bool single, bool first, bool perm, bool narrow, int logr, const PassArgs<FP> &a,
                     uint32_t grid, hipStream_t s);

// The lazy 32-bit field (FieldU32L) has FieldU32's argument layout: the
// same words, tables and constants.
inline PassArgs<FieldU32L> lazy_args(const PassArgs<FieldU32> &a) {
    static_assert(sizeof(PassArgs<FieldU32L>) == sizeof(PassArgs<FieldU32>), "copy every PassArgs field below");
    PassArgs<FieldU32L> b;
    b.in = a.in;
    b.out = a.out;
    b.hi = a.hi;
    b.lo = a.lo;
    b.log_n = a.log_n;
    b.log_j = a.log_j;
    b.log_rmax = a.log_rmax;
    b.batches = a.batches;
    b.tiles_per_transform = a.tiles_per_transform;
    b.flags = a.flags;
    b.rt = a.rt;
    b.mul_in = a.mul_in;
    b.in_tab = a.in_tab;
    b.out_tab = a.out_tab;
    return b;
}

// 32-bit transforms with p < 2^30 run the lazy kernels.  STNTT_U32_LAZY=0 in
// the environment keeps every modulus on the reduced ones (same words; for
// A/B timing).
inline bool u32_use_lazy(uint32_t p) {
    static const bool enabled = [] {
        const char *e = getenv("STNTT_U32_LAZY");
        return !(e && e[0] == '0');
    }();
    return enabled && u32_lazy_modulus(p);
}

template <class FP>
inline void launch_pass(int dir, bool single, bool first, bool perm, bool narrow, int logr, const PassArgs<FP> &a,
                        uint32_t grid, hipStream_t s) {
    if constexpr (std::is_same_v<FP, FieldU32>) {
        if (u32_use_lazy(a.rt.p)) {
            const PassArgs<FieldU32L> b = lazy_args(a);
            if (dir == 0) launch_pass_dir<FieldU32L, 0>(single, first, perm, narrow, logr, b, grid, s);
            else launch_pass_dir<FieldU32L, 1>(single, first, perm, narrow, logr, b, grid, s);
            return;
        }
    }
    if (dir == 0) launch_pass_dir<FP, 0>(single, first, perm, narrow, logr, a, grid, s);
    else launch_pass_dir<FP, 1>(single, first, perm, narrow, logr, a, grid, s);
}

// One pass of the large-radix kernel (ntt_big.hpp): radix 2^logr (10..12),
// 2^logc columns per tile (at most big_full_logc, at least two fewer),
// `ntiles` tiles over the whole batch, a persistent grid of at most
// num_cus times the kernel's occupancy workgroups.
template <class FP, int DIR>
void launch_big_dir(bool first, int logr, int logc, const PassArgs<FP> &a, uint32_t ntiles, int num_cus,
                    hipStream_t s);

template <class FP>
inline void launch_big(int dir, bool first, int logr, int logc, const PassArgs<FP> &a, uint32_t ntiles, int num_cus,
                       hipStream_t s) {
    if constexpr (std::is_same_v<FP, FieldU32>) {
        if (u32_use_lazy(a.rt.p)) {
            const PassArgs<FieldU32L> b = lazy_args(a);
            if (dir == 0) launch_big_dir<FieldU32L, 0>(first, logr, logc, b, ntiles, num_cus, s);
            else launch_big_dir<FieldU32L, 1>(first, logr, logc, b, ntiles, num_cus, s);
            return;
        }
    }
    if (dir == 0) launch_big_dir<FP, 0>(first, logr, logc, a, ntiles, num_cus, s);
    else launch_big_dir<FP, 1>(first, logr, logc, a, ntiles, num_cus, s);
}

// The persistent descriptor-queue kernel (ntt_batch.hpp) and its counter
// reset; occupancy of the former in workgroups per CU.
void launch_batch_reset(uint32_t *counter, hipStream_t s);
void launch_batch(uint32_t grid, const BatchTask *tasks, uint32_t count, uint32_t *counter, uint64_t *scratch,
                  uint64_t scratch_stride, hipStream_t s);
hipError_t batch_occupancy(int *per_cu);

}  // namespace stntt
