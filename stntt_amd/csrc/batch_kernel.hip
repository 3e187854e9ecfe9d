// batch_kernel.hip -- launchers of the persistent descriptor-queue kernel
// (ntt_batch.hpp), in a translation unit of their own.
#include "launch.hpp"

namespace stntt {

// Persistent kernel: grid = resident workgroups; each pulls task indices
// from *counter until they run out.  scratch holds one slice of
// scratch_stride elements (of the widest field) per workgroup.
__global__ void __launch_bounds__(BATCH_THREADS) ntt_batch_kernel(const BatchTask *tasks, uint32_t count,
                                                                  uint32_t *counter, uint64_t *scratch,
                                                                  uint64_t scratch_stride) {
    __shared__ __align__(16) unsigned char smem[BATCH_LDS_BYTES];
    __shared__ uint32_t next;
    uint64_t *my_scratch = scratch + (uint64_t)blockIdx.x * scratch_stride;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
        __syncthreads();
        // uniform: the descriptor loads into scalar registers, once (a
        // reference into global memory is re-read after every store)
        const uint32_t idx = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (idx >= count) break;
        const BatchTask t = tasks[idx];
        if (t.field == 1) {
            if (t.dir == 0) batch_run_task<FieldGL, 0>(t, my_scratch, smem);
            else batch_run_task<FieldGL, 1>(t, my_scratch, smem);
        } else {
            if (t.dir == 0) batch_run_task<FieldU32, 0>(t, (uint32_t *)my_scratch, smem);
            else batch_run_task<FieldU32, 1>(t, (uint32_t *)my_scratch, smem);
        }
    }
}

// Zeroes the queue counter (a kernel, so graph replays need no memset node).
__global__ void ntt_batch_reset(uint32_t *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

void launch_batch_reset(uint32_t *counter, hipStream_t s) {
    hipLaunchKernelGGL(ntt_batch_reset, dim3(1), dim3(64), 0, s, counter);
}

void launch_batch(uint32_t grid, const BatchTask *tasks, uint32_t count, uint32_t *counter, uint64_t *scratch,
                  uint64_t scratch_stride, hipStream_t s) {
    hipLaunchKernelGGL(ntt_batch_kernel, dim3(grid), dim3(BATCH_THREADS), 0, s, tasks, count, counter, scratch,
                       scratch_stride);
}

hipError_t batch_occupancy(int *per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ntt_batch_kernel, BATCH_THREADS, 0);
}

}  // namespace stntt
