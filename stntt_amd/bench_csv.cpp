// bench_csv.cpp -- the reference's benchmark driver (src/bench.cu:9-165)
// over the C++ drop-in header, printing the same CSV so the reference's
// post-processing (metrics/stats.py: median per log2_N and Block size)
// reads it unchanged.
//
// The same shape as the reference: a BenchStats record holding the
// cuda::ntt::NttArgs of the run, runNttIteration<I> timing findParams ->
// NttArgs -> autoNtt<n> on batch 1 of the reference's input (mt19937 seed
// 42, ints 0..9, src/util/rng.h:10,33-35), runAll doing 10 warm-up runs per
// size and then `reps` timed ones, for each maximum block size {64, 128,
// 256, 512, 1024}.  Columns, all in milliseconds as the reference's Timer
// (src/util/timer.h): the kernel time autoNtt returns (HIP events around
// the launches), findParams, NttArgs construction, autoNtt including its
// host<->device copies, and the whole iteration.  "Block size" is read from
// args.dimBlock as the reference's toCsv does (src/bench.cu:153-165): the
// reference's launch geometry for that n and maximum; this library accepts
// the value and picks its own tiles, and the column keeps the rows aligned
// with the reference's.
//
// Usage: bench_csv [reps]   (default 5; sizes 2^2..2^20 as the reference)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#include "../include/stntt/ntt_kernel.hpp"

using Clock = std::chrono::high_resolution_clock;
using Ms = std::chrono::duration<double, std::milli>;

struct Timer {
    Clock::time_point t0 = Clock::now();
    void start() { t0 = Clock::now(); }
    Ms stop() const { return Clock::now() - t0; }
};

struct BenchStats {
    cuda::ntt::NttArgs args;
    Ms gpuTime{};
    Ms findParamsTime{};
    Ms initArgsTime{};
    Ms kernelWithInitTime{};
    Ms totalTime{};
};

template <size_t I>
BenchStats runNttIteration(std::span<int> vec, unsigned maxBlockSize) {
    Timer totalTimer;
    BenchStats stats;

    constexpr size_t n = 1UL << I;
    size_t batches = 1;

    Timer timer;
    auto [root, mod] = findParams(n, 11);
    stats.findParamsTime = timer.stop();

    timer.start();
    cuda::ntt::NttArgs args(vec.subspan(0, n * batches), n, batches, root, mod, 2, maxBlockSize);
    stats.args = args;
    stats.initArgsTime = timer.stop();

    timer.start();
    double gpuTime = cuda::ntt::autoNtt<n>(args);
    stats.kernelWithInitTime = timer.stop();
    stats.gpuTime = Ms(gpuTime);

    stats.totalTime = totalTimer.stop();
    return stats;
}

template <size_t... I>
std::vector<BenchStats> runAll(std::span<int> vec, size_t reps, unsigned maxBlockSize, std::index_sequence<I...>) {
    std::vector<BenchStats> result;
    result.reserve(sizeof...(I) * reps);
    // warm-up
    ([&]() {
        for (size_t r = 0; r < 10; r++) runNttIteration<I + 2>(vec, maxBlockSize);
    }(), ...);
    ([&]() {
        for (size_t r = 0; r < reps; r++) result.push_back(runNttIteration<I + 2>(vec, maxBlockSize));
    }(), ...);
    return result;
}

int main(int argc, char **argv) {
    const size_t reps = argc > 1 ? (size_t)std::atoi(argv[1]) : 5;
    if (reps < 1 || reps > 1000) {
        std::fprintf(stderr, "usage: bench_csv [reps in 1..1000]\n");
        return 2;
    }
    std::mt19937 engine(42);
    std::uniform_int_distribution<int> dist{0, 9};
    std::vector<int> vec((size_t)1 << 20);
    for (auto &e : vec) e = dist(engine);

    std::vector<BenchStats> stats;
    for (unsigned maxBlockSize : cuda::ntt::defaultMaxBlockSizes) {
        auto blockStats = runAll(vec, reps, maxBlockSize, std::make_index_sequence<19>{});  // 2^2..2^20
        stats.insert(stats.end(), blockStats.begin(), blockStats.end());
    }
    std::printf("log2_N,Block size,Kernel time,findParams time,initArgs time,Kernel+transfers time,Total time\n");
    for (const BenchStats &p : stats)
        std::printf("%u,%u,%.6f,%.6f,%.6f,%.6f,%.6f\n", cuda::log2_uint(p.args.n),
                    p.args.dimBlock.x * p.args.dimBlock.y * p.args.dimBlock.z, p.gpuTime.count(),
                    p.findParamsTime.count(), p.initArgsTime.count(), p.kernelWithInitTime.count(),
                    p.totalTime.count());
    return 0;
}
