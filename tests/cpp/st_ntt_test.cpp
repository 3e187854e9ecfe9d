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
    // nttCt / inttGs / pwcNtt / pwc_convolution as the reference calls them
    // (src/ntt/ntt_cpu.cpp:11-42, 130-192), against the oracle's restatements
    for (size_t L : {3u, 10u, 15u}) {
        size_t n = 1ull << L;
        auto [root, mod] = stntt::findParams(n, 11);
        std::vector<int> x(vec.begin(), vec.begin() + n), y(vec.begin() + n, vec.begin() + 2 * n);
        std::vector<uint32_t> ox(x.begin(), x.end()), oy(y.begin(), y.end());
        std::vector<int> ct = x;
        stntt::nttCt(ct, (int)root, (int)mod);
        orc_ntt_ct_u32(ox.data(), n, (uint32_t)root, (uint32_t)mod);
        if (!std::equal(ct.begin(), ct.end(), ox.begin())) { std::printf("nttCt mismatch 2^%zu\n", L); ++failures; }
        stntt::inttGs(ct, (int)root, (int)mod);
        if (ct != x) { std::printf("inttGs round trip mismatch 2^%zu\n", L); ++failures; }
        std::vector<uint32_t> pa(x.begin(), x.end()), pb(y.begin(), y.end());
        orc_pwc_ntt_u32(pa.data(), pb.data(), n, (uint32_t)root, (uint32_t)mod);
        std::vector<int> a = x, b = y;
        stntt::pwcNtt(a, b, (int)root, (int)mod);
        if (!std::equal(a.begin(), a.end(), pa.begin())) { std::printf("pwcNtt mismatch 2^%zu\n", L); ++failures; }
        cases += 3;
    }
    {
        std::vector<int> s0 = {1, 2, 3, 4}, s1 = {5, 6, 7, 8};
        auto [res, root, mod] = stntt::pwc_convolution(s0, s1);
        std::vector<int> want(4, 0);  // schoolbook cyclic product
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) want[(i + j) % 4] += s0[i] * s1[j];
        if (res != want) { std::printf("pwc_convolution mismatch\n"); ++failures; }
        ++cases;
    }
    std::printf("%d cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
