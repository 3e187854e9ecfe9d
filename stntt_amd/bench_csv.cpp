// bench_csv.cpp -- the reference's benchmark driver (src/bench.cu:26-165)
// replayed over the C++ drop-in header, printing the same CSV so the
// reference's post-processing (metrics/stats.py: median per log2_N and
// Block size) reads it unchanged.
//
// For each maximum block size {64, 128, 256, 512, 1024} and log2 n = 2..20:
// 10 warm-up runs, then 5 timed runs of findParams -> NttArgs -> autoNtt<n>
// on batch 1 of the reference's input (mt19937 seed 42, ints 0..9,
// src/util/rng.h:10,33-35).  Columns, all in milliseconds as the
// reference's Timer (src/util/timer.h): the kernel time autoNtt returns
// (HIP events around the launches), findParams, NttArgs construction,
// autoNtt including its host<->device copies, and the whole iteration.
// "Block size" is the reference's launch geometry for that n and maximum
// (getNttKernelArgs, src/cuda/ntt/ntt_kernel.cuh:31-39); this library
// accepts the value and picks its own tiles, and the column is kept so the
// rows line up with the reference's.
//
// Usage: bench_csv [min_log2 max_log2]   (defaults 2 20)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../include/stntt/ntt_kernel.hpp"

using Clock = std::chrono::high_resolution_clock;

static double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct Row {
    unsigned log2n, block;
    double kernel, find_params, init_args, kernel_with_init, total;
};

static Row run_once(std::vector<int> &vec, unsigned log2n, unsigned max_block) {
    auto t_total = Clock::now();
    Row r{};
    r.log2n = log2n;
    const unsigned n = 1u << log2n, batches = 1;
    auto t = Clock::now();
    auto [root, mod] = stntt::findParams(n, 11);
    r.find_params = ms_since(t);
    t = Clock::now();
    stntt::NttArgs args(std::span<int>(vec.data(), (size_t)n * batches), n, batches, (unsigned)root, (unsigned)mod, 2,
                        max_block);
    r.init_args = ms_since(t);
    // dimBlock.x * dimBlock.y of the reference's geometry for radix 2
    const unsigned half = n / 2;
    r.block = std::min(half, max_block) * std::max(std::min(max_block / half, batches), 1u);
    t = Clock::now();
    r.kernel = stntt::autoNtt(args);
    r.kernel_with_init = ms_since(t);
    r.total = ms_since(t_total);
    return r;
}

int main(int argc, char **argv) {
    unsigned lo = 2, hi = 20;
    if (argc == 3) {
        lo = (unsigned)std::atoi(argv[1]);
        hi = (unsigned)std::atoi(argv[2]);
    }
    if (lo < 1 || hi > 27 || lo > hi) {
        std::fprintf(stderr, "usage: bench_csv [min_log2 max_log2] with 1 <= min <= max <= 27\n");
        return 2;
    }
    std::mt19937 engine(42);
    std::uniform_int_distribution<int> dist{0, 9};
    std::vector<int> vec((size_t)1 << hi);
    for (auto &e : vec) e = dist(engine);

    std::vector<Row> rows;
    for (unsigned max_block : stntt::defaultMaxBlockSizes) {
        for (unsigned L = lo; L <= hi; ++L)
            for (int r = 0; r < 10; ++r) run_once(vec, L, max_block);  // warm-up
        for (unsigned L = lo; L <= hi; ++L)
            for (int r = 0; r < 5; ++r) rows.push_back(run_once(vec, L, max_block));
    }
    std::printf("log2_N,Block size,Kernel time,findParams time,initArgs time,Kernel+transfers time,Total time\n");
    for (const Row &r : rows)
        std::printf("%u,%u,%.6f,%.6f,%.6f,%.6f,%.6f\n", r.log2n, r.block, r.kernel, r.find_params, r.init_args,
                    r.kernel_with_init, r.total);
    return 0;
}
