// st_ntt_test.cpp -- the reference's GPU test (src/tests/cuda/ntt/st_ntt_test.cu:16-97)
// replayed through the C++ drop-in header with the reference's own calls:
// unqualified findParams(n, 11), nttStockham(cpuVec, n, root, mod, batches),
// cuda::ntt::NttArgs(gpuVec, n, batches, root, mod, 2, maxBlockSize),
// cuda::ntt::autoNtt<n>(args), then cpuVec == gpuVec; same inputs (values
// 0..9, src/util/rng.h:10,33-35), batches {1,2,5,10}, block sizes {64..1024},
// log2 n = 2..20.  The drop-in keeps the src/ntt entry points, so here
// nttStockham runs on the GPU too; every iteration therefore also compares
// with the oracle (oracle/ntt_oracle.c), which equals the reference's CPU
// code wherever its int arithmetic does not wrap (log2 n <= 13).  Then the
// reference's other entry points as it declares them: nttStockham with
// nsteps (both overloads), nttStockhamIdx, nttCt, inttGs, pwcNtt,
// pwc_convolution, and the error kinds.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#include "../../include/stntt/ntt_kernel.hpp"
#include "../../oracle/ntt_oracle.h"

static int failures = 0;

#define CHECK_MESSAGE(cond, ...)                    \
    do {                                            \
        if (!(cond)) {                              \
            std::printf("FAILED: " __VA_ARGS__);    \
            std::printf("\n");                      \
            ++failures;                             \
        }                                           \
    } while (0)

template <size_t I>
void runNttIteration(std::span<int> vec, size_t batches, unsigned maxBlockSize) {
    constexpr size_t n = 1UL << I;

    auto [root, mod] = findParams(n, 11);

    auto refSpan = vec.subspan(0, n * batches);

    std::vector<int> cpuVec{refSpan.begin(), refSpan.end()};
    nttStockham(cpuVec, n, root, mod, batches);

    std::vector<int> gpuVec{refSpan.begin(), refSpan.end()};
    cuda::ntt::NttArgs args(gpuVec, n, batches, root, mod, 2, maxBlockSize);

    cuda::ntt::autoNtt<n>(args);
    CHECK_MESSAGE(cpuVec == gpuVec, "N: 2^%zu | NUM_BATCHES: %zu | BLOCK_SIZE: %u", I, batches, args.blockSize);

    std::vector<int> orcVec{refSpan.begin(), refSpan.end()};
    orc_stockham_u32_batch(reinterpret_cast<uint32_t *>(orcVec.data()), n, (uint32_t)root, (uint32_t)mod, batches, -1);
    CHECK_MESSAGE(orcVec == gpuVec, "oracle N: 2^%zu | NUM_BATCHES: %zu", I, batches);
}

template <size_t... I>
void runAll(std::span<int> vec, size_t batches, unsigned maxBlockSize, std::index_sequence<I...>) {
    ([&]() { runNttIteration<I + 2>(vec, batches, maxBlockSize); }(), ...);
}

int main() {
    std::mt19937 engine(42);  // src/util/rng.h:10
    std::uniform_int_distribution<int> dist{0, 9};
    std::vector<int> vec((1 << 20) * 10);
    for (auto &e : vec) e = dist(engine);
    int cases = 0;
    for (unsigned batch : cuda::ntt::defaultBatchesNums)
        for (unsigned bs : cuda::ntt::defaultMaxBlockSizes) {
            runAll(vec, batch, bs, std::make_index_sequence<19>{});  // log2 n = 2..20
            cases += 19;
        }

    // nsteps: the first k stages (both overloads), against the oracle's stage loop
    for (int L : {1, 4, 9, 13}) {
        const int n = 1 << L;
        auto [root, mod] = findParams(n, 11);
        for (int k : {0, 1, L / 2, L, L + 2}) {
            std::vector<int> v(vec.begin(), vec.begin() + 3 * n);
            std::vector<int> o = v;
            nttStockham(v, n, (int)root, (int)mod, 3, k);
            orc_stockham_u32_batch(reinterpret_cast<uint32_t *>(o.data()), n, (uint32_t)root, (uint32_t)mod, 3, k);
            CHECK_MESSAGE(v == o, "nsteps %d at 2^%d", k, L);
            std::vector<int> w(vec.begin(), vec.begin() + n);
            nttStockham(w.data(), n, (int)root, (int)mod, k);
            CHECK_MESSAGE(std::equal(w.begin(), w.end(), o.begin()), "int* nsteps %d at 2^%d", k, L);
            cases += 2;
        }
        // nttStockhamIdx: with every stage, the bit-reversal permutation
        std::vector<int> idx(2 * n);
        for (int i = 0; i < 2 * n; ++i) idx[i] = i;
        nttStockhamIdx(idx, n, (int)root, (int)mod, 2);
        bool brev = true;
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < n; ++i) brev = brev && idx[b * n + (int)bitReverse(i, L)] == b * n + i;
        CHECK_MESSAGE(brev, "nttStockhamIdx is not the bit reversal at 2^%d", L);
        std::vector<int> part(n), opart(n);
        for (int i = 0; i < n; ++i) part[i] = opart[i] = 7 * i + 1;
        nttStockhamIdx(part.data(), n, (int)root, (int)mod, L / 2);
        orc_stockham_idx_mod(opart.data(), n, (uint32_t)mod, L / 2);
        CHECK_MESSAGE(part == opart, "nttStockhamIdx nsteps %d at 2^%d", L / 2, L);
        cases += 2;
    }

    // error behaviour of the reference API
    bool threw = false;
    try {
        std::vector<int> v(12);
        nttStockham(v, 12, 3, 13, 1);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK_MESSAGE(threw, "no invalid_argument for n = 12");
    threw = false;
    try {
        std::vector<int> v(8);
        cuda::ntt::autoNtt(cuda::ntt::NttArgs(v, 8, 2, 9, 17, 2));
    } catch (const std::out_of_range &) {
        threw = true;
    }
    CHECK_MESSAGE(threw, "no out_of_range for a short span");

    // Goldilocks round trip
    std::vector<uint64_t> g(1 << 16), h;
    std::mt19937_64 e64(7);
    for (auto &x : g) x = e64() % 0xFFFFFFFF00000001ull;
    h = g;
    stntt::autoNttGoldilocks(h, 1 << 16, 1);
    std::vector<uint64_t> o = g;
    orc_stockham_gl(o.data(), 16, -1);
    CHECK_MESSAGE(o == h, "Goldilocks forward mismatch");
    stntt::autoNttGoldilocks(h, 1 << 16, 1, true);
    CHECK_MESSAGE(g == h, "Goldilocks round trip mismatch");

    // nttCt / inttGs / pwcNtt / pwc_convolution as the reference calls them
    // (src/ntt/ntt_cpu.cpp:11-42, 130-192), against the oracle's restatements
    for (size_t L : {3u, 10u, 15u}) {
        size_t n = 1ull << L;
        auto [root, mod] = findParams(n, 11);
        std::vector<int> x(vec.begin(), vec.begin() + n), y(vec.begin() + n, vec.begin() + 2 * n);
        std::vector<uint32_t> ox(x.begin(), x.end());
        std::vector<int> ct = x;
        nttCt(ct, (int)root, (int)mod);
        orc_ntt_ct_u32(ox.data(), n, (uint32_t)root, (uint32_t)mod);
        CHECK_MESSAGE(std::equal(ct.begin(), ct.end(), ox.begin()), "nttCt mismatch 2^%zu", L);
        inttGs(ct, (int)root, (int)mod);
        CHECK_MESSAGE(ct == x, "inttGs round trip mismatch 2^%zu", L);
        std::vector<uint32_t> pa(x.begin(), x.end()), pb(y.begin(), y.end());
        orc_pwc_ntt_u32(pa.data(), pb.data(), n, (uint32_t)root, (uint32_t)mod);
        std::vector<int> a = x, b = y;
        pwcNtt(a, b, (int)root, (int)mod);
        CHECK_MESSAGE(std::equal(a.begin(), a.end(), pa.begin()), "pwcNtt mismatch 2^%zu", L);
        cases += 3;
    }
    {
        std::vector<int> s0 = {1, 2, 3, 4}, s1 = {5, 6, 7, 8};
        auto [res, root, mod] = pwc_convolution(s0, s1);
        std::vector<int> want(4, 0);  // schoolbook cyclic product
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) want[(i + j) % 4] += s0[i] * s1[j];
        CHECK_MESSAGE(res == want, "pwc_convolution mismatch");
        ++cases;
    }
    std::printf("%d cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
