// launch_pass.hip -- the pass kernels of one field and one direction.
//
// Compiled once per (field, direction), with STNTT_TU_FIELD (FieldGL,
// FieldU32 or FieldM64) and STNTT_TU_DIR (0 forward, 1 inverse) defined on
// the command line, so that the template instances of the library compile
// in parallel translation units.
#include "launch.hpp"
#include "plan.hpp"

#if !defined(STNTT_TU_FIELD) || !defined(STNTT_TU_DIR)
#error "compile with -DSTNTT_TU_FIELD=<FieldGL|FieldU32|FieldM64> -DSTNTT_TU_DIR=<0|1>"
#endif

namespace stntt {

template <class FP, int DIR, int LOGR>
static void launch_single(const PassArgs<FP> &a, uint32_t grid, bool narrow, hipStream_t s) {
    constexpr int LOGC = pass_logc<FP>(LOGR, true);
    if constexpr (LOGR <= NARROW_MAX_LOGR) {
        if (narrow) {  // one transform, one wave
            hipLaunchKernelGGL((ntt_pass_kernel<FP, LOGR, 0, NARROW_THREADS, DIR, true>), dim3(grid),
                               dim3(NARROW_THREADS), 0, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((ntt_pass_kernel<FP, LOGR, LOGC, PASS_THREADS, DIR, true>), dim3(grid), dim3(PASS_THREADS), 0,
                       s, a);
}

template <class FP, int DIR, int LOGR>
static void launch_tile(bool first, bool perm, const PassArgs<FP> &a, uint32_t grid, hipStream_t s) {
    constexpr int NT = tile_threads<FP, LOGR>();
    // perm: bit-reversed input on the first pass, bit-reversed output on the last
    if (first) {
        if (perm) hipLaunchKernelGGL((ntt_tile_kernel<FP, LOGR, DIR, true, PERM_IN>), dim3(grid), dim3(NT), 0, s, a);
        else hipLaunchKernelGGL((ntt_tile_kernel<FP, LOGR, DIR, true>), dim3(grid), dim3(NT), 0, s, a);
    } else {
        if (perm) hipLaunchKernelGGL((ntt_tile_kernel<FP, LOGR, DIR, false, PERM_OUT>), dim3(grid), dim3(NT), 0, s, a);
        else hipLaunchKernelGGL((ntt_tile_kernel<FP, LOGR, DIR, false>), dim3(grid), dim3(NT), 0, s, a);
    }
}

template <class FP, int DIR, int... L>
static void launch_single_impl(int logr, const PassArgs<FP> &a, uint32_t grid, bool narrow, hipStream_t s,
                               std::integer_sequence<int, L...>) {
    ((logr == L + 1 ? (launch_single<FP, DIR, L + 1>(a, grid, narrow, s), 0) : 0), ...);
}

template <class FP, int DIR, int... L>
static void launch_tile_impl(int logr, bool first, bool perm, const PassArgs<FP> &a, uint32_t grid, hipStream_t s,
                             std::integer_sequence<int, L...>) {
    ((logr == L + MIN_LOGR_TILE ? (launch_tile<FP, DIR, L + MIN_LOGR_TILE>(first, perm, a, grid, s), 0) : 0), ...);
}

template <class FP, int DIR>
void launch_pass_dir(bool single, bool first, bool perm, bool narrow, int logr, const PassArgs<FP> &a,
                     uint32_t grid, hipStream_t s) {
    using SeqS = std::make_integer_sequence<int, MAX_LOGR_SINGLE<FP>>;
    using SeqT = std::make_integer_sequence<int, MAX_LOGR_TILE - MIN_LOGR_TILE + 1>;
    if (single) launch_single_impl<FP, DIR>(logr, a, grid, narrow, s, SeqS{});
    else launch_tile_impl<FP, DIR>(logr, first, perm, a, grid, s, SeqT{});
}

template void launch_pass_dir<STNTT_TU_FIELD, STNTT_TU_DIR>(bool, bool, bool, bool, int,
                                                             const PassArgs<STNTT_TU_FIELD> &, uint32_t, hipStream_t);

}  // namespace stntt
