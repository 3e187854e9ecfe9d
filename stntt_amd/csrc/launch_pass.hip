// launch_pass.hip -- the pass kernels of one field and one direction.
//
// Compiled once per (field, direction), with STNTT_TU_FIELD (FieldGL,
// FieldU32, FieldU32L or FieldM64) and STNTT_TU_DIR (0 forward, 1 inverse) defined on
// the command line, so that the template instances of the library compile
// in parallel translation units.
#include "launch.hpp"
#include "plan.hpp"

#if !defined(STNTT_TU_FIELD) || !defined(STNTT_TU_DIR)
#error "compile with -DSTNTT_TU_FIELD=<FieldGL|FieldU32|FieldU32L|FieldM64> -DSTNTT_TU_DIR=<0|1>"
#endif

namespace stntt {

template <class FP, int DIR, int LOGR>
static void launch_single(const PassArgs<FP> &a, uint32_t grid, bool narrow, hipStream_t s) {
    constexpr int LOGC = pass_logc<FP>(LOGR, true);
    if constexpr (LOGR <= narrow_max_logr<FP>()) {
        if (narrow) {  // one transform per workgroup
            constexpr int NT = narrow_threads<FP, LOGR>();
            hipLaunchKernelGGL((ntt_pass_kernel<FP, LOGR, 0, NT, DIR, true>), dim3(grid), dim3(NT), 0, s, a);
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

// the large-radix kernel for one (LOGR, LOGC): a persistent grid of at most
// num_cus x occupancy workgroups, a multiple of 8 so each stays on its XCD
template <class FP, int DIR, int LOGR, int LOGC>
static void launch_big_one(bool first, const PassArgs<FP> &a, uint32_t ntiles, int num_cus, hipStream_t s) {
    constexpr int NT = big_threads<FP, LOGR, LOGC>();
    static int occ[2] = {0, 0};  // workgroups per CU, first pass / later passes
    int &o = occ[first ? 0 : 1];
    if (!o) {
        int v = 0;
        if (first) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, ntt_big_kernel<FP, LOGR, LOGC, DIR, true>, NT, 0);
        else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, ntt_big_kernel<FP, LOGR, LOGC, DIR, false>, NT, 0);
        o = v > 0 ? v : 1;
    }
    uint64_t grid = (uint64_t)(num_cus > 0 ? num_cus : 1) * o;
    if (grid > ntiles) grid = ntiles;
    if (grid >= 8) grid -= grid % 8;
    if (first) hipLaunchKernelGGL((ntt_big_kernel<FP, LOGR, LOGC, DIR, true>), dim3((uint32_t)grid), dim3(NT), 0, s, a, ntiles);
    else hipLaunchKernelGGL((ntt_big_kernel<FP, LOGR, LOGC, DIR, false>), dim3((uint32_t)grid), dim3(NT), 0, s, a, ntiles);
}

template <class FP, int DIR, int LOGR>
static void launch_big_r(bool first, int logc, const PassArgs<FP> &a, uint32_t ntiles, int num_cus, hipStream_t s) {
    constexpr int F = big_full_logc<FP>(LOGR);
    if (logc == F) launch_big_one<FP, DIR, LOGR, F>(first, a, ntiles, num_cus, s);
    else if (logc == F - 1) launch_big_one<FP, DIR, LOGR, F - 1>(first, a, ntiles, num_cus, s);
    else if (logc == F - 2) launch_big_one<FP, DIR, LOGR, F - 2>(first, a, ntiles, num_cus, s);
}

template <class FP, int DIR>
void launch_big_dir(bool first, int logr, int logc, const PassArgs<FP> &a, uint32_t ntiles, int num_cus,
                    hipStream_t s) {
    if (logr == 10) launch_big_r<FP, DIR, 10>(first, logc, a, ntiles, num_cus, s);
    else if (logr == 11) launch_big_r<FP, DIR, 11>(first, logc, a, ntiles, num_cus, s);
    else if (logr == 12) launch_big_r<FP, DIR, 12>(first, logc, a, ntiles, num_cus, s);
}

template void launch_pass_dir<STNTT_TU_FIELD, STNTT_TU_DIR>(bool, bool, bool, bool, int,
                                                             const PassArgs<STNTT_TU_FIELD> &, uint32_t, hipStream_t);
template void launch_big_dir<STNTT_TU_FIELD, STNTT_TU_DIR>(bool, int, int, const PassArgs<STNTT_TU_FIELD> &, uint32_t,
                                                            int, hipStream_t);

}  // namespace stntt
