// stntt/ntt_kernel.hpp -- C++ drop-in for the reference's GPU host API and
// CPU entry points, implemented over the C ABI of libstntt.so (stntt.h).
//
// Declares, with the reference's names, scopes, signatures, argument
// meaning and exceptions:
//   global scope, as src/ntt/ntt_cpu.h:7-20 and src/ntt/ntt_util.h:11-96:
//     findParams, findModulus, findPrimitiveRoot, findGenerator,
//     isPrimitiveRoot, uniquePrimeFactors, isPrime, sqrt, pow, modulo,
//     elemWiseMul, bitReverse, bitReverseVector,
//     nttStockham (int* and span overloads, std::optional<int> nsteps),
//     nttStockhamIdx (both overloads), nttCt, inttGs (defined at
//     src/ntt/ntt_cpu.cpp:130, not declared by the reference), pwcNtt,
//     pwc_convolution;
//   namespace cuda::ntt, as src/cuda/ntt/ntt_kernel.cuh:24-82,152-201:
//     defaultBlockSize, defaultBatchesNums, defaultMaxBlockSizes,
//     KernelArgs, getNttKernelArgs, NttArgs (with dimGrid, dimBlock,
//     sharedMem, isGlobal and initKernelArgs()), autoNtt<n>;
//   namespace cuda, as src/cuda/math.cuh:10-35: log2_uint, log2,
//     log2_constexpr.
// A caller switches by including this header in place of
// "cuda/ntt/ntt_kernel.cuh", "ntt/ntt_cpu.h" and "ntt/ntt_util.h" and
// linking libstntt.so (and libamdhip64).
//
// Where the work runs: every transform entry point runs on the GPU
// (nttStockham and autoNtt on the multi-pass kernels; nsteps and
// nttStockhamIdx on the reference's radix-2 stage loop, a device launch per
// stage).  The parameter search and the small host helpers stay on the
// host, as in the reference.  Arithmetic is exact in Z_p: where the
// reference's int products wrap (p >= 46341, i.e. n >= 2^14 with
// findParams(n, 11)) the results are the field's, not the wrapped ints.
// The launch geometry in NttArgs is the reference's formula, kept for
// callers that read it (src/bench.cu:128-165); the library picks its own
// tiles.
// Errors: std::invalid_argument (bad size or parameters), std::out_of_range
// (host span shorter than n * batches), std::runtime_error (device).
#pragma once

#include <algorithm>
#include <array>
#include <bit>
#include <concepts>
#include <cstdint>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../stntt.h"

namespace stntt {

inline void check(int rc) {
    if (rc == STNTT_OK) return;
    std::string msg = stntt_last_error();
    if (rc == STNTT_EINVAL) throw std::invalid_argument(msg);
    if (rc == STNTT_ERANGE) throw std::out_of_range(msg);
    throw std::runtime_error("stntt error " + std::to_string(rc) + ": " + msg);
}

// RAII owner of a stntt_ctx (stream, resident twiddles, scratch).
class Context {
public:
    explicit Context(int device = 0) { check(stntt_ctx_create(device, &ctx_)); }
    ~Context() { stntt_ctx_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    stntt_ctx *get() const { return ctx_; }

private:
    stntt_ctx *ctx_ = nullptr;
};

// One context per process for the reference-shaped free functions (the
// reference keeps its twiddles in process-global device symbols).
inline Context &defaultContext() {
    static Context ctx(0);
    return ctx;
}

// three unsigned extents, the shape of CUDA's dim3 (only .x .y .z are read
// by the reference's callers)
struct Dim3 {
    unsigned x = 1, y = 1, z = 1;
    constexpr Dim3() = default;
    constexpr Dim3(unsigned x_, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

// a mod m for the reference's signed operands, in [0, m)
inline uint64_t reduce_signed(int64_t a, uint64_t m) {
    int64_t r = a % (int64_t)m;
    return (uint64_t)(r < 0 ? r + (int64_t)m : r);
}

__extension__ typedef unsigned __int128 wide_t;  // exact products of 64-bit words

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((wide_t)a * b % m); }

inline uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    for (b %= m; e; e >>= 1, b = mulmod(b, b, m))
        if (e & 1) r = mulmod(r, b, m);
    return r;
}

}  // namespace stntt

// ---------------------------------------------------------------------------
// src/ntt/ntt_util.h
// ---------------------------------------------------------------------------

// (root, mod): the smallest prime mod = k size + 1 >= min_mod, root =
// g^((mod-1)/size) with g its smallest generator (ntt_util.cpp:10-14).
inline std::tuple<size_t, size_t> findParams(size_t size, size_t min_mod) {
    uint64_t root = 0, mod = 0;
    stntt::check(stntt_find_params(size, min_mod, &root, &mod));
    return {root, mod};
}

// the smallest prime k veclength + 1 >= min (ntt_util.cpp:16-27)
inline size_t findModulus(size_t veclength, size_t min) {
    uint64_t root = 0, mod = 0;
    stntt::check(stntt_find_params(veclength, min, &root, &mod));
    return mod;
}

// floor(sqrt(n)) (ntt_util.cpp:72-84)
inline size_t sqrt(size_t n) {
    size_t r = 0;
    for (size_t bit = (size_t)1 << (std::bit_width(n) / 2 + 1); bit; bit >>= 1)
        if ((stntt::wide_t)(r + bit) * (r + bit) <= n) r += bit;
    return r;
}

// trial division by 2..floor(sqrt(n)), as ntt_util.h:65-69 (which makes
// 0 and 1 "prime": their divisor range is empty)
template <typename T>
inline bool isPrime(T n) {
    for (T x = 2; (stntt::wide_t)x * x <= (stntt::wide_t)n; ++x)
        if (n % x == 0) return false;
    return true;
}

// distinct prime factors of n, ascending (ntt_util.cpp:51-69)
inline std::vector<size_t> uniquePrimeFactors(size_t n) {
    std::vector<size_t> f;
    for (size_t d = 2; d <= n / d; ++d)
        if (n % d == 0) {
            f.push_back(d);
            while (n % d == 0) n /= d;
        }
    if (n > 1) f.push_back(n);
    return f;
}

// val^degree == 1 and val^(degree/q) != 1 for every prime q | degree
// (ntt_util.cpp:44-49)
inline bool isPrimitiveRoot(size_t val, size_t degree, size_t mod) {
    if (stntt::powmod(val, degree, mod) != 1) return false;
    for (size_t q : uniquePrimeFactors(degree))
        if (stntt::powmod(val, degree / q, mod) == 1) return false;
    return true;
}

// the smallest generator of Z_mod^* (totient = mod - 1; ntt_util.cpp:36-42)
inline size_t findGenerator(size_t totient, size_t mod) {
    for (size_t g = 1; g < mod; ++g)
        if (isPrimitiveRoot(g, totient, mod)) return g;
    throw std::runtime_error("No generator exists");
}

// g^(totient / degree) for the smallest generator g (ntt_util.cpp:29-34)
inline size_t findPrimitiveRoot(size_t degree, size_t totient, size_t mod) {
    return stntt::powmod(findGenerator(totient, mod), totient / degree, mod);
}

// x mod m in [0, m) for signed x (ntt_util.cpp:88-94)
inline int modulo(int x, int mod) {
    int r = x % mod;
    return r >= 0 ? r : r + mod;
}

// modular power; a negative exponent takes the inverse by Fermat's little
// theorem, as ntt_util.h:72-96 (exact: products in 128 bits)
template <std::integral I>
I pow(I base, I exponent, I mod) {
    const uint64_t m = (uint64_t)mod;
    uint64_t b = std::is_signed_v<I> ? stntt::reduce_signed((int64_t)base, m) : (uint64_t)base % m;
    if constexpr (std::is_signed_v<I>) {
        if (exponent < 0) {
            b = stntt::powmod(b, m - 2, m);
            exponent = -exponent;
        }
    }
    return (I)stntt::powmod(b, (uint64_t)exponent, m);
}

// s0[i] = s0[i] s1[i] mod p for i < s0.size() <= s1.size() (ntt_util.h:28-33).
// Signed operands are first brought into [0, mod) as the reference's
// modulo() brings its product there; the product is then exact (the
// reference forms it in T and wraps for |s0[i] s1[i]| >= 2^31).
template <typename T>
void elemWiseMul(std::span<T> s0, std::span<T> s1, size_t mod) {
    if (s1.size() < s0.size()) throw std::out_of_range("s1 is shorter than s0");
    auto red = [mod](T v) -> uint64_t {
        if constexpr (std::is_signed_v<T>) return stntt::reduce_signed((int64_t)v, mod);
        else return (uint64_t)v % mod;
    };
    for (size_t i = 0; i < s0.size(); ++i) s0[i] = (T)stntt::mulmod(red(s0[i]), red(s1[i]), mod);
}

// the low `bits` bits of x reversed (ntt_util.h:35-44)
template <typename T>
size_t bitReverse(T x, size_t bits) {
    size_t y = 0;
    for (size_t i = 0; i < bits; ++i, x >>= 1) y = (y << 1) | (size_t)(x & 1);
    return y;
}

// element i moves to bitReverse(i, levels), in place (ntt_util.h:46-63)
template <typename T>
std::vector<T> &bitReverseVector(std::vector<T> &vec, size_t levels) {
    for (size_t i = 0; i < vec.size(); ++i) {
        const size_t j = bitReverse(i, levels);
        if (j > i) std::swap(vec[i], vec[j]);
    }
    return vec;
}

template <typename T>
std::vector<T> &bitReverseVector(std::vector<T> &vec) {
    return bitReverseVector(vec, std::bit_width(vec.size()) - 1);
}

// ---------------------------------------------------------------------------
// src/cuda/math.cuh (host forms)
// ---------------------------------------------------------------------------
namespace cuda {
inline unsigned int log2_uint(unsigned int x) { return 31 - __builtin_clz(x); }
inline unsigned log2(unsigned x) {
    unsigned r = 0;
    while (x >>= 1) ++r;
    return r;
}
constexpr unsigned log2_constexpr(unsigned x) { return x <= 1 ? 0 : 1 + log2_constexpr(x >> 1); }
}  // namespace cuda

// ---------------------------------------------------------------------------
// src/cuda/ntt/ntt_kernel.cuh
// ---------------------------------------------------------------------------
namespace cuda::ntt {

using dim3 = stntt::Dim3;
constexpr unsigned defaultBlockSize = 1024;                                            // ntt_kernel.cuh:25
constexpr std::array<unsigned, 4> defaultBatchesNums = {1, 2, 5, 10};                  // ntt_kernel.cuh:26
constexpr std::array<unsigned, 5> defaultMaxBlockSizes = {64, 128, 256, 512, 1024};    // ntt_kernel.cuh:27
constexpr unsigned warpSizeConst = 32;                                                  // cu_util.cuh:9

using KernelArgs = std::tuple<dim3, dim3, unsigned>;

// The reference's launch geometry (ntt_kernel.cuh:30-39): grid, block and
// shared bytes its kernels would use for this (n, radix, batches, block
// size).  Informational here.
inline KernelArgs getNttKernelArgs(unsigned n, unsigned radix, unsigned batches,
                                   unsigned blockSize = defaultBlockSize) {
    const unsigned per = radix ? n / radix : 0;
    dim3 grid{blockSize ? (per * batches + blockSize - 1) / blockSize : 0};
    dim3 block{std::min(per, blockSize), std::max(std::min(per ? blockSize / per : 0u, batches), 1u)};
    const unsigned stepsInWarp = cuda::log2(warpSizeConst * radix);
    const unsigned sharedMem =
        (cuda::log2(n) > stepsInWarp) ? std::min(n / 2, blockSize) * block.y * (unsigned)sizeof(int) : 0;
    return {grid, block, sharedMem};
}

// Same fields, constructors and initKernelArgs() as the reference
// (ntt_kernel.cuh:41-82), plus `inverse` (the reference has no inverse
// GPU path).
struct NttArgs {
    unsigned n{};
    unsigned batches{};
    unsigned root{};
    unsigned mod{};
    unsigned radix{};

    unsigned blockSize{};
    dim3 dimGrid{};
    dim3 dimBlock{};
    unsigned sharedMem{};
    // True if the reference would need global memory and double buffering
    bool isGlobal{};
    bool inverse{};

    std::span<int> vec;

    NttArgs() = default;
    NttArgs(unsigned n, unsigned batches, unsigned root, unsigned mod, unsigned radix,
            unsigned blockSize = defaultBlockSize)
        : n(n), batches(batches), root(root), mod(mod), radix(radix), blockSize(blockSize) {
        initKernelArgs();
        if (radix && n / radix > dimBlock.x) isGlobal = true;
    }
    NttArgs(std::span<int> vec, unsigned n, unsigned batches, unsigned root, unsigned mod, unsigned radix,
            unsigned blockSize = defaultBlockSize)
        : NttArgs(n, batches, root, mod, radix, blockSize) {
        this->vec = vec;
    }

    void initKernelArgs() { std::tie(dimGrid, dimBlock, sharedMem) = getNttKernelArgs(n, radix, batches, blockSize); }
};

// autoNtt with a runtime size: copies args.vec[0, n*batches) to the device,
// transforms in place, copies back; returns kernel-only milliseconds.
inline float autoNtt(const NttArgs &args, stntt::Context &ctx = stntt::defaultContext()) {
    float ms = 0.f;
    stntt::check(stntt_auto_ntt_i32(ctx.get(), args.vec.data(), args.vec.size(), args.n, args.batches, args.root,
                                    args.mod, args.inverse ? STNTT_INVERSE : STNTT_FORWARD, &ms));
    return ms;
}

// The reference's compile-time form (ntt_kernel.cuh:152-201); the template
// size must equal args.n (the reference does not check this; here a
// mismatch is invalid_argument).
template <size_t n>
float autoNtt(NttArgs args) {
    if (args.n != n) throw std::invalid_argument("autoNtt<n>: template size differs from args.n");
    return autoNtt(args, stntt::defaultContext());
}

}  // namespace cuda::ntt

// ---------------------------------------------------------------------------
// src/ntt/ntt_cpu.h (run on the GPU)
// ---------------------------------------------------------------------------
namespace stntt {
inline uint32_t root_word(int root, int mod) { return (uint32_t)reduce_signed(root, (uint64_t)mod); }
inline void positive_size(int size) {
    if (size <= 0) throw std::invalid_argument("size (" + std::to_string(size) + ") must be a power of 2");
}
}  // namespace stntt

// nttStockham(span, size, root, mod, batch, nsteps) (ntt_cpu.cpp:87-90):
// `batch` rows of `size` ints transformed in place.  Without nsteps, the
// whole forward transform on the multi-pass path; with nsteps, the first
// nsteps radix-2 stages of the reference's loop (a device launch per stage,
// the reference's strides and ping-pong; 0 or negative: no stage).
inline void nttStockham(std::span<int> vec, int size, int root, int mod, int batch,
                        std::optional<int> nsteps = std::nullopt) {
    stntt::positive_size(size);
    if (batch <= 0) return;  // no rows (the reference's loop runs zero times)
    float ms = 0.f;
    if (!nsteps) {
        cuda::ntt::autoNtt(cuda::ntt::NttArgs(vec, (unsigned)size, (unsigned)batch, (unsigned)root, (unsigned)mod, 2));
        return;
    }
    stntt::check(stntt_auto_stockham_steps_i32(stntt::defaultContext().get(), vec.data(), vec.size(),
                                               (uint32_t)size, (uint32_t)batch, stntt::root_word(root, mod),
                                               (uint32_t)mod, *nsteps, &ms));
}

// nttStockham(int*, size, root, mod, nsteps) (ntt_cpu.cpp:46-85): one
// transform of the `size` ints at vec
inline void nttStockham(int *vec, int size, int root, int mod, std::optional<int> nsteps = std::nullopt) {
    stntt::positive_size(size);
    nttStockham(std::span<int>(vec, (size_t)size), size, root, mod, 1, nsteps);
}

// nttStockhamIdx (ntt_cpu.cpp:92-127): the Stockham data movement without
// arithmetic, on the device (with every stage, the bit-reversal
// permutation).  root is unused, as in the reference; mod only sets the
// stage strides pow(2, i, mod).
inline void nttStockhamIdx(std::span<int> vec, int size, int root, int mod, int batches,
                           std::optional<int> nsteps = std::nullopt) {
    (void)root;
    stntt::positive_size(size);
    if (batches <= 0) return;
    const int steps = nsteps ? *nsteps : (int)std::bit_width((unsigned)size) - 1;
    float ms = 0.f;
    stntt::check(stntt_auto_stockham_idx_i32(stntt::defaultContext().get(), vec.data(), vec.size(), (uint32_t)size,
                                             (uint32_t)batches, (uint32_t)mod, steps, &ms));
}

inline void nttStockhamIdx(int *vec, int size, int root, int mod, std::optional<int> nsteps = std::nullopt) {
    stntt::positive_size(size);
    nttStockhamIdx(std::span<int>(vec, (size_t)size), size, root, mod, 1, nsteps);
}

// nttCt (ntt_cpu.cpp:11-42): natural order in, the transform in
// bit-reversed order out, in place.
inline void nttCt(std::vector<int> &vec, int root, int mod) {
    float ms = 0.f;
    stntt::check(stntt_auto_ntt_i32_ex(stntt::defaultContext().get(), vec.data(), vec.size(), (uint32_t)vec.size(), 1,
                                       stntt::root_word(root, mod), (uint32_t)mod, STNTT_FORWARD, STNTT_BITREV_OUT,
                                       &ms));
}

// inttGs (ntt_cpu.cpp:130-164): bit-reversed order in, the inverse (times
// n^-1) in natural order out, in place; root is the forward root, as in the
// reference.
inline void inttGs(std::vector<int> &vec, int root, int mod) {
    float ms = 0.f;
    stntt::check(stntt_auto_ntt_i32_ex(stntt::defaultContext().get(), vec.data(), vec.size(), (uint32_t)vec.size(), 1,
                                       stntt::root_word(root, mod), (uint32_t)mod, STNTT_INVERSE, STNTT_BITREV_IN,
                                       &ms));
}

// pwcNtt (ntt_cpu.cpp:167-174): vec0 <- cyclic convolution of vec0 and
// vec1, vec1 <- nttCt(vec1).
inline void pwcNtt(std::vector<int> &vec0, std::vector<int> &vec1, int root, int mod) {
    if (vec0.size() != vec1.size()) throw std::invalid_argument("vec0 and vec1 differ in size");
    float ms = 0.f;
    stntt::check(stntt_auto_polymul_i32(stntt::defaultContext().get(), vec0.data(), vec1.data(), (uint32_t)vec0.size(),
                                        stntt::root_word(root, mod), (uint32_t)mod, &ms));
}

// pwc_convolution (ntt_cpu.cpp:180-192): the cyclic convolution of two
// positive vectors of equal power-of-two length, with (root, mod) =
// findParams(len, max^2 len + 1) so no output wraps.  The reference forms
// max^2 len + 1 in int; here it is exact and a modulus of 2^31 or more is
// invalid_argument.
inline std::tuple<std::vector<int>, int, int> pwc_convolution(std::span<const int> s0, std::span<const int> s1) {
    if (s0.empty() || s0.size() != s1.size()) throw std::invalid_argument("s0 and s1 must be non-empty, equal length");
    int64_t max_val = std::max(*std::max_element(s0.begin(), s0.end()), *std::max_element(s1.begin(), s1.end()));
    uint64_t min_mod = (uint64_t)(max_val * max_val) * s0.size() + 1;
    if (min_mod >= (1ull << 31)) throw std::invalid_argument("max^2 * len + 1 needs a modulus of 2^31 or more");
    auto [root, mod] = findParams(s0.size(), min_mod);
    std::vector<int> vec0(s0.begin(), s0.end()), vec1(s1.begin(), s1.end());
    pwcNtt(vec0, vec1, (int)root, (int)mod);
    return {vec0, (int)root, (int)mod};
}

// ---------------------------------------------------------------------------
// beyond the reference
// ---------------------------------------------------------------------------
namespace stntt {

// the inverse of nttStockham: natural order in and out, times n^-1
inline void inttStockham(std::span<int> vec, int size, int root, int mod, int batches) {
    positive_size(size);
    cuda::ntt::NttArgs a(vec, (unsigned)size, (unsigned)batches, root_word(root, mod), (unsigned)mod, 2);
    a.inverse = true;
    cuda::ntt::autoNtt(a);
}

// Goldilocks (P = 2^64 - 2^32 + 1), which the reference declares but does
// not implement (src/cuda/ntt/implementations/goldilocks/arithmetic.cuh).
inline float autoNttGoldilocks(std::span<uint64_t> vec, unsigned n, unsigned batches, bool inverse = false,
                               Context &ctx = defaultContext()) {
    float ms = 0.f;
    check(stntt_auto_ntt_goldilocks(ctx.get(), vec.data(), vec.size(), n, batches,
                                    inverse ? STNTT_INVERSE : STNTT_FORWARD, &ms));
    return ms;
}

// the reference-shaped names, also reachable as stntt::name
using ::elemWiseMul;
using ::findParams;
using ::inttGs;
using ::nttCt;
using ::nttStockham;
using ::nttStockhamIdx;
using ::pwc_convolution;
using ::pwcNtt;
using cuda::ntt::autoNtt;
using cuda::ntt::defaultBatchesNums;
using cuda::ntt::defaultBlockSize;
using cuda::ntt::defaultMaxBlockSizes;
using cuda::ntt::NttArgs;

}  // namespace stntt
