// stntt/ntt_kernel.hpp -- C++ drop-in for the reference's GPU host API and
// CPU entry points, implemented over the C ABI of libstntt.so (stntt.h).
//
// Replaces, with the same names, argument meaning and exceptions:
//   cuda::ntt::NttArgs            src/cuda/ntt/ntt_kernel.cuh:41-82
//   cuda::ntt::autoNtt<n>(args)   src/cuda/ntt/ntt_kernel.cuh:152-201
//   findParams(size, min_mod)     src/ntt/ntt_util.cpp:10-14
//   nttStockham(span, size, root, mod, batches)   src/ntt/ntt_cpu.cpp:87-90
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
