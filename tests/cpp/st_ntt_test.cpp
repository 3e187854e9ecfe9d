// st_ntt_test.cpp -- the reference's GPU test (src/tests/cuda/ntt/st_ntt_test.cu:16-97)
// replayed through the C++ drop-in header: same call sequence
// (findParams -> CPU nttStockham -> NttArgs -> autoNtt<n> -> compare), same
// inputs (values 0..9), batches {1,2,5,10} and block sizes {64..1024}.
// The CPU side is the oracle (oracle/ntt_oracle.c); for log2 n <= 13, where
// the reference's int arithmetic does not overflow, it equals the reference.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#include "../../include/stntt/ntt_kernel.hpp"
#include "../../oracle/ntt_oracle.h"

static int failures = 0;

template <size_t I>
static void runNttIteration(std::span<int> vec, size_t batches, unsigned maxBlockSize) {
    constexpr size_t n = 1UL << I;
    auto [root, mod] = stntt::findParams(n, 11);
    auto refSpan = vec.subspan(0, n * batches);
    std::vector<int> cpuVec{refSpan.begin(), refSpan.end()};
    if (orc_stockham_u32_batch(reinterpret_cast<uint32_t *>(cpuVec.data()), n, (uint32_t)root, (uint32_t)mod,
                               batches, -1) != 0) {
        std::printf("oracle failed\n");
        ++failures;
        return;
    }
    std::vector<int> gpuVec{refSpan.begin(), refSpan.end()};
    cuda::ntt::NttArgs args(gpuVec, n, batches, root, mod, 2, maxBlockSize);
    cuda::ntt::autoNtt<n>(args);
    if (cpuVec != gpuVec) {
        std::printf("MISMATCH N: 2^%zu | NUM_BATCHES: %zu | BLOCK_SIZE: %u\n", I, batches, args.blockSize);
        ++failures;
    }
}

template <size_t... I>
static void runAll(std::span<int> vec, size_t batches, unsigned maxBlockSize, std::index_sequence<I...>) {
    (runNttIteration<I + 2>(vec, batches, maxBlockSize), ...);
}

int main() {
    std::mt19937 engine(42);  // src/util/rng.h:10
    std::uniform_int_distribution<int> dist{0, 9};
    std::vector<int> vec((1 << 20) * 10);
    for (auto &e : vec) e = dist(engine);
    int cases = 0;
    for (unsigned batch : stntt::defaultBatchesNums)
        for (unsigned bs : stntt::defaultMaxBlockSizes) {
            runAll(vec, batch, bs, std::make_index_sequence<19>{});  // log2 n = 2..20
            cases += 19;
        }
    // error behaviour of the reference API
    bool threw = false;
    try {
        std::vector<int> v(12);
        stntt::nttStockham(v, 12, 3, 13, 1);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    if (!threw) { std::printf("no invalid_argument for n = 12\n"); ++failures; }
    threw = false;
    try {
        std::vector<int> v(8);
        cuda::ntt::autoNtt(cuda::ntt::NttArgs(v, 8, 2, 9, 17, 2));
    } catch (const std::out_of_range &) {
        threw = true;
    }
    if (!threw) { std::printf("no out_of_range for a short span\n"); ++failures; }
    // Goldilocks round trip
    std::vector<uint64_t> g(1 << 16), h;
    std::mt19937_64 e64(7);
    for (auto &x : g) x = e64() % 0xFFFFFFFF00000001ull;
    h = g;
    stntt::autoNttGoldilocks(h, 1 << 16, 1);
    std::vector<uint64_t> o = g;
    orc_stockham_gl(o.data(), 16, -1);
    if (o != h) { std::printf("Goldilocks forward mismatch\n"); ++failures; }
    stntt::autoNttGoldilocks(h, 1 << 16, 1, true);
    if (g != h) { std::printf("Goldilocks round trip mismatch\n"); ++failures; }
    std::printf("%d cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
