// ref_shim.cpp -- C entry points over the reference's own CPU NTT sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's
// src/ntt/ntt_cpu.cpp and src/ntt/ntt_util.cpp (read in place, never copied)
// into oracle/_ref/libstntt_ref.so by oracle/Makefile.  Used to pin the
// oracle and to write tests/golden fixtures; the product never loads it.
#include "ntt/ntt_cpu.h"
#include "ntt/ntt_util.h"

#include <cstdint>
#include <cstring>
#include <exception>
#include <optional>
#include <span>
#include <vector>

// Defined at src/ntt/ntt_cpu.cpp:130 but absent from src/ntt/ntt_cpu.h.
void inttGs(std::vector<int> &vec, int root, int mod);

namespace {
std::optional<int> steps_opt(int nsteps) {
    return nsteps < 0 ? std::nullopt : std::optional<int>(nsteps);
}
}  // namespace

extern "C" {

// findParams(size, min_mod) -> (root, mod): src/ntt/ntt_util.cpp:10-14
int ref_find_params(uint64_t n, uint64_t min_mod, uint64_t *root, uint64_t *mod) {
    try {
        auto [r, m] = findParams(n, min_mod);
        *root = r;
        *mod = m;
        return 0;
    } catch (const std::exception &) {
        return -2;
    }
}

// nttStockham(span, size, root, mod, batches, nsteps): src/ntt/ntt_cpu.cpp:87-90
int ref_ntt_stockham(int32_t *vec, int size, int root, int mod, int batches, int nsteps) {
    try {
        nttStockham(std::span<int>(vec, (size_t)size * batches), size, root, mod, batches, steps_opt(nsteps));
        return 0;
    } catch (const std::invalid_argument &) {
        return -1;
    }
}

// nttStockhamIdx(span, ...): src/ntt/ntt_cpu.cpp:124-127.  mod only feeds
// J = pow(2, i, mod) (:105), so it must exceed n: a 31-bit prime is used.
int ref_ntt_stockham_idx(int32_t *vec, int size, int batches, int nsteps) {
    try {
        nttStockhamIdx(std::span<int>(vec, (size_t)size * batches), size, 0, 2147483647, batches, steps_opt(nsteps));
        return 0;
    } catch (const std::invalid_argument &) {
        return -1;
    }
}

// nttStockhamIdx with the caller's modulus, which sets the stage strides
// J = pow(2, i, mod) (src/ntt/ntt_cpu.cpp:105)
int ref_ntt_stockham_idx_mod(int32_t *vec, int size, int batches, int mod, int nsteps) {
    try {
        nttStockhamIdx(std::span<int>(vec, (size_t)size * batches), size, 0, mod, batches, steps_opt(nsteps));
        return 0;
    } catch (const std::invalid_argument &) {
        return -1;
    }
}

// elemWiseMul<int>(s0, s1, mod): src/ntt/ntt_util.h:28-33
int ref_elem_wise_mul(int32_t *s0, int32_t *s1, int size, uint64_t mod) {
    elemWiseMul(std::span<int>(s0, (size_t)size), std::span<int>(s1, (size_t)size), (size_t)mod);
    return 0;
}

// nttCt(vector&, root, mod): src/ntt/ntt_cpu.cpp:11-42
int ref_ntt_ct(int32_t *vec, int size, int root, int mod) {
    std::vector<int> v(vec, vec + size);
    try {
        nttCt(v, root, mod);
    } catch (const std::invalid_argument &) {
        return -1;
    }
    std::memcpy(vec, v.data(), sizeof(int) * (size_t)size);
    return 0;
}

// inttGs(vector&, root, mod): src/ntt/ntt_cpu.cpp:130-164
int ref_intt_gs(int32_t *vec, int size, int root, int mod) {
    std::vector<int> v(vec, vec + size);
    try {
        inttGs(v, root, mod);
    } catch (const std::invalid_argument &) {
        return -1;
    }
    std::memcpy(vec, v.data(), sizeof(int) * (size_t)size);
    return 0;
}

// pwc_convolution(s0, s1) -> (result, root, mod): src/ntt/ntt_cpu.cpp:180-192
int ref_pwc_convolution(const int32_t *a, const int32_t *b, int size, int32_t *out, int *root, int *mod) {
    try {
        auto [res, r, m] = pwc_convolution(std::span<const int>(a, size), std::span<const int>(b, size));
        std::memcpy(out, res.data(), sizeof(int) * (size_t)size);
        *root = r;
        *mod = m;
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}
}
