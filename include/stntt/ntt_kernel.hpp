// stntt/ntt_kernel.hpp -- C++ drop-in for the reference's GPU host API and
// CPU entry points, implemented over the C ABI of libstntt.so (stntt.h).
//
// Replaces, with the same names, argument meaning and exceptions:
//   cuda::ntt::NttArgs            src/cuda/ntt/ntt_kernel.cuh:41-82
//   cuda::ntt::autoNtt<n>(args)   src/cuda/ntt/ntt_kernel.cuh:152-201
//   findParams(size, min_mod)     src/ntt/ntt_util.cpp:10-14
//   nttStockham(span, size, root, mod, batches)   src/ntt/ntt_cpu.cpp:87-90
//   nttCt(vec, root, mod)         src/ntt/ntt_cpu.cpp:11-42
//   inttGs(vec, root, mod)        src/ntt/ntt_cpu.cpp:130-164
//   pwcNtt(vec0, vec1, root, mod) src/ntt/ntt_cpu.cpp:167-174
//   pwc_convolution(s0, s1)       src/ntt/ntt_cpu.cpp:180-192
//   elemWiseMul(s0, s1, mod)      src/ntt/ntt_util.h:28-33
// A caller switches by including this header instead of
// "cuda/ntt/ntt_kernel.cuh" / "ntt/ntt_cpu.h" and linking libstntt.so.
// Errors: std::invalid_argument (bad size or parameters), std::out_of_range
// (host span shorter than n * batches), std::runtime_error (device).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <tuple>
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

inline std::tuple<size_t, size_t> findParams(size_t size, size_t min_mod) {
    uint64_t root = 0, mod = 0;
    check(stntt_find_params(size, min_mod, &root, &mod));
    return {root, mod};
}

constexpr unsigned defaultBlockSize = 1024;                                    // ntt_kernel.cuh:25
constexpr std::array<unsigned, 4> defaultBatchesNums = {1, 2, 5, 10};          // ntt_kernel.cuh:26
constexpr std::array<unsigned, 5> defaultMaxBlockSizes = {64, 128, 256, 512, 1024};  // ntt_kernel.cuh:27

// Same fields and constructors as the reference.  radix and blockSize are
// kept for call compatibility; the MI355X path picks its own tiles.
struct NttArgs {
    unsigned n{};
    unsigned batches{};
    unsigned root{};
    unsigned mod{};
    unsigned radix{};
    unsigned blockSize{};
    bool isGlobal{};
    bool inverse{};
    std::span<int> vec;

    NttArgs() = default;
    NttArgs(unsigned n, unsigned batches, unsigned root, unsigned mod, unsigned radix,
            unsigned blockSize = defaultBlockSize)
        : n(n), batches(batches), root(root), mod(mod), radix(radix), blockSize(blockSize) {
        unsigned half = radix ? n / radix : n;
        isGlobal = half > std::min(half, blockSize);
    }
    NttArgs(std::span<int> vec, unsigned n, unsigned batches, unsigned root, unsigned mod, unsigned radix,
            unsigned blockSize = defaultBlockSize)
        : NttArgs(n, batches, root, mod, radix, blockSize) {
        this->vec = vec;
    }
};

// autoNtt with a runtime size: copies args.vec[0, n*batches) to the device,
// transforms in place, copies back; returns kernel-only milliseconds.
inline float autoNtt(const NttArgs &args, Context &ctx = defaultContext()) {
    float ms = 0.f;
    check(stntt_auto_ntt_i32(ctx.get(), args.vec.data(), args.vec.size(), args.n, args.batches, args.root,
                             args.mod, args.inverse ? STNTT_INVERSE : STNTT_FORWARD, &ms));
    return ms;
}

// The reference's compile-time form; the template size must match args.n
// (the reference does not check this; here a mismatch is invalid_argument).
template <size_t n>
float autoNtt(const NttArgs &args) {
    if (args.n != n) throw std::invalid_argument("autoNtt<n>: template size differs from args.n");
    return autoNtt(args);
}

// nttStockham(span, size, root, mod, batches) on the GPU, in place.
inline void nttStockham(std::span<int> vec, int size, int root, int mod, int batches) {
    if (size <= 0) throw std::invalid_argument("size must be positive");
    autoNtt(NttArgs(vec, (unsigned)size, (unsigned)batches, (unsigned)root, (unsigned)mod, 2));
}

inline void inttStockham(std::span<int> vec, int size, int root, int mod, int batches) {
    if (size <= 0) throw std::invalid_argument("size must be positive");
    NttArgs a(vec, (unsigned)size, (unsigned)batches, (unsigned)root, (unsigned)mod, 2);
    a.inverse = true;
    autoNtt(a);
}

// nttCt: natural order in, the transform in bit-reversed order out, in place.
inline void nttCt(std::vector<int> &vec, int root, int mod) {
    float ms = 0.f;
    check(stntt_auto_ntt_i32_ex(defaultContext().get(), vec.data(), vec.size(), (uint32_t)vec.size(), 1,
                                (uint32_t)root, (uint32_t)mod, STNTT_FORWARD, STNTT_BITREV_OUT, &ms));
}

// inttGs: bit-reversed order in, the inverse (times n^-1) in natural order
// out, in place; root is the forward root, as in the reference.
inline void inttGs(std::vector<int> &vec, int root, int mod) {
    float ms = 0.f;
    check(stntt_auto_ntt_i32_ex(defaultContext().get(), vec.data(), vec.size(), (uint32_t)vec.size(), 1,
                                (uint32_t)root, (uint32_t)mod, STNTT_INVERSE, STNTT_BITREV_IN, &ms));
}

// pwcNtt: vec0 <- cyclic convolution of vec0 and vec1, vec1 <- nttCt(vec1).
inline void pwcNtt(std::vector<int> &vec0, std::vector<int> &vec1, int root, int mod) {
    if (vec0.size() != vec1.size()) throw std::invalid_argument("vec0 and vec1 differ in size");
    float ms = 0.f;
    check(stntt_auto_polymul_i32(defaultContext().get(), vec0.data(), vec1.data(), (uint32_t)vec0.size(),
                                 (uint32_t)root, (uint32_t)mod, &ms));
}

// elemWiseMul: s0[i] = s0[i] * s1[i] mod p for i < s0.size() <= s1.size(),
// values in [0, mod).  Host-side, like the reference (it is one pass of
// multiplies; the fused device form is inside pwcNtt).
template <typename T>
void elemWiseMul(std::span<T> s0, std::span<T> s1, size_t mod) {
    if (s1.size() < s0.size()) throw std::out_of_range("s1 is shorter than s0");
    __extension__ typedef unsigned __int128 wide_t;  // exact products of 64-bit words
    for (size_t i = 0; i < s0.size(); ++i) s0[i] = (T)(((wide_t)(uint64_t)s0[i] * (uint64_t)s1[i]) % mod);
}

// pwc_convolution: the cyclic convolution of two positive vectors of equal
// power-of-two length, with (root, mod) = findParams(len, max^2 len + 1) so
// no output wraps.  The reference forms max^2 len + 1 in int; here it is
// exact and a modulus of 2^31 or more is invalid_argument.
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

// Goldilocks (P = 2^64 - 2^32 + 1), which the reference declares but does
// not implement (src/cuda/ntt/implementations/goldilocks/arithmetic.cuh).
inline float autoNttGoldilocks(std::span<uint64_t> vec, unsigned n, unsigned batches, bool inverse = false,
                               Context &ctx = defaultContext()) {
    float ms = 0.f;
    check(stntt_auto_ntt_goldilocks(ctx.get(), vec.data(), vec.size(), n, batches,
                                    inverse ? STNTT_INVERSE : STNTT_FORWARD, &ms));
    return ms;
}

}  // namespace stntt

// Source compatibility with code written against the reference.
namespace cuda::ntt {
using stntt::autoNtt;
using stntt::defaultBatchesNums;
using stntt::defaultBlockSize;
using stntt::defaultMaxBlockSizes;
using stntt::NttArgs;
}  // namespace cuda::ntt
