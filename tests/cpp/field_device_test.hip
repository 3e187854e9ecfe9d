// field_device_test.hip -- the field arithmetic of stntt_amd/csrc/field_*.hpp
// run ON THE GPU on edge and random values, checked against 128-bit integer
// arithmetic on the host.  The kernels' paths that random transform data
// almost never reaches (a borrow in the Goldilocks reduction needs
// lo < hi >> 32, a carry out of a Montgomery sum needs p > 2^63 and large
// operands) are fed directly here.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#include "../../stntt_amd/csrc/field_gl.hpp"
#include "../../stntt_amd/csrc/field_m64.hpp"
#include "../../stntt_amd/csrc/field_u32.hpp"

using namespace stntt;
typedef unsigned __int128 u128;

#define HIP_CHECK(x)                                                                   \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            std::exit(2);                                                              \
        }                                                                              \
    } while (0)

constexpr int NPOW = 192;  // every shift amount of gl_mul_pow2 / gl_sub_mul_pow2

// Goldilocks: out[0..n) mul, [n..2n) reduce128, [2n..3n) add, [3n..4n) sub
__global__ void k_gl(const uint64_t *a, const uint64_t *b, const uint64_t *ca, const uint64_t *cb, uint64_t *out,
                     uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = gl_mul(a[i], b[i]);
    out[n + i] = gl_reduce128(a[i], b[i]);
    out[2 * n + i] = gl_add(ca[i], cb[i]);
    out[3 * n + i] = gl_sub(ca[i], cb[i]);
}

template <int S>
__device__ void pow2_one(const uint64_t *x, const uint64_t *ca, const uint64_t *cb, uint64_t *out, uint32_t n,
                         uint32_t i) {
    out[(uint64_t)S * n + i] = gl_mul_pow2<S>(x[i]);
    out[(uint64_t)(NPOW + S) * n + i] = gl_sub_mul_pow2<S>(ca[i], cb[i]);
}

template <int... S>
__device__ void pow2_all(const uint64_t *x, const uint64_t *ca, const uint64_t *cb, uint64_t *out, uint32_t n,
                         uint32_t i, std::integer_sequence<int, S...>) {
    (pow2_one<S>(x, ca, cb, out, n, i), ...);
}

// out[S n + i] = x 2^S, out[(192 + S) n + i] = (a - b) 2^S
__global__ void k_pow2(const uint64_t *x, const uint64_t *ca, const uint64_t *cb, uint64_t *out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pow2_all(x, ca, cb, out, n, i, std::make_integer_sequence<int, NPOW>{});
}

// 64-bit Montgomery field for one prime: mul (any b), add, sub (canonical)
__global__ void k_m64(const uint64_t *a, const uint64_t *b, const uint64_t *ca, const uint64_t *cb, uint64_t *out,
                      uint32_t n, uint64_t p, uint64_t pinv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = m64_mul(ca[i], b[i], p, pinv);
    out[n + i] = m64_add(ca[i], cb[i], p);
    out[2 * n + i] = m64_sub(ca[i], cb[i], p);
    (void)a;
}

// 32-bit field for one prime: Montgomery product with any 32-bit word, add, sub
__global__ void k_u32(const uint32_t *x, const uint32_t *wm, const uint32_t *ca, const uint32_t *cb, uint32_t *out,
                      uint32_t n, uint32_t p, uint32_t pinv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = u32_mont(x[i], wm[i], p, pinv);
    out[n + i] = u32_add(ca[i], cb[i], p);
    out[2 * n + i] = u32_sub(ca[i], cb[i], p);
}

// 32-bit field, lazy form: operands la, lb in [0, 2p), wm < p
__global__ void k_u32l(const uint32_t *la, const uint32_t *lb, const uint32_t *wm, uint32_t *out, uint32_t n,
                       uint32_t p, uint32_t pinv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = u32l_add(la[i], lb[i], 2 * p);
    out[n + i] = u32l_sub(la[i], lb[i], 2 * p);
    out[2 * n + i] = u32l_sub_mont(la[i], lb[i], wm[i], p, pinv);
    out[3 * n + i] = u32l_mont(la[i], lb[i], p, pinv);
    out[4 * n + i] = u32l_canon(la[i], p);
}

static int failures = 0;
static long cases = 0;
#define CHECK(cond, ...)                        \
    do {                                        \
        ++cases;                                \
        if (!(cond)) {                          \
            if (failures < 20) {                \
                std::printf(__VA_ARGS__);       \
                std::printf("\n");              \
            }                                   \
            ++failures;                         \
        }                                       \
    } while (0)

template <class T>
static T *dev_copy(const std::vector<T> &v) {
    T *d = nullptr;
    HIP_CHECK(hipMalloc(&d, v.size() * sizeof(T)));
    HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

template <class T>
static std::vector<T> host_copy(const T *d, size_t count) {
    std::vector<T> v(count);
    HIP_CHECK(hipMemcpy(v.data(), d, count * sizeof(T), hipMemcpyDeviceToHost));
    return v;
}

static uint64_t gl_ref_pow2(uint64_t x, int s) {
    u128 r = x % GL_P;
    for (int i = 0; i < s; ++i) r = (r * 2) % GL_P;
    return (uint64_t)r;
}

int main() {
    std::mt19937_64 rng(20261204);
    // words that stress carries and borrows: near 0, 2^32, P, 2^63 and 2^64
    std::vector<uint64_t> edge = {0,
                                  1,
                                  2,
                                  GL_EPS - 1,
                                  GL_EPS,
                                  GL_EPS + 1,
                                  1ull << 32,
                                  (1ull << 32) + 1,
                                  GL_P - 2,
                                  GL_P - 1,
                                  GL_P,
                                  GL_P + 1,
                                  ~0ull,
                                  ~0ull - 1,
                                  1ull << 63,
                                  (1ull << 63) - 1,
                                  0xFFFFFFFF00000000ull,
                                  0x00000000FFFFFFFEull,
                                  0x8000000080000000ull,
                                  0x00000001FFFFFFFFull};
    std::vector<uint64_t> a, b;
    for (uint64_t x : edge)
        for (uint64_t y : edge) {
            a.push_back(x);
            b.push_back(y);
        }
    // reduce128(lo, hi) with lo < hi >> 32 (the borrow of lo - h1), with and
    // without a carry of h0 EPS + s, and random words
    for (int i = 0; i < 4000; ++i) {
        const uint64_t h1 = rng() | (1ull << 31), h0 = rng() & 0xFFFFFFFFull;
        a.push_back(rng() % h1);
        b.push_back((h1 << 32) | h0);
    }
    for (int i = 0; i < 20000; ++i) {
        a.push_back(rng());
        b.push_back(rng());
    }
    const uint32_t n = (uint32_t)a.size();
    std::vector<uint64_t> ca(n), cb(n);
    for (uint32_t i = 0; i < n; ++i) {
        ca[i] = a[i] % GL_P;
        cb[i] = b[i] % GL_P;
    }
    const dim3 block(256), grid((n + 255) / 256);

    // ---- Goldilocks: mul, reduce128, add, sub ----
    {
        uint64_t *da = dev_copy(a), *db = dev_copy(b), *dca = dev_copy(ca), *dcb = dev_copy(cb), *dout = nullptr;
        HIP_CHECK(hipMalloc(&dout, 4ull * n * sizeof(uint64_t)));
        hipLaunchKernelGGL(k_gl, grid, block, 0, 0, da, db, dca, dcb, dout, n);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        const auto out = host_copy(dout, 4ull * n);
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t mul = (uint64_t)((u128)a[i] * b[i] % GL_P);
            const uint64_t red = (uint64_t)((((u128)b[i] << 64) | a[i]) % GL_P);
            CHECK(out[i] == mul, "gl_mul %016llx %016llx", (unsigned long long)a[i], (unsigned long long)b[i]);
            CHECK(out[n + i] == red, "gl_reduce128 lo %016llx hi %016llx", (unsigned long long)a[i],
                  (unsigned long long)b[i]);
            CHECK(out[2 * n + i] == (uint64_t)(((u128)ca[i] + cb[i]) % GL_P), "gl_add");
            CHECK(out[3 * n + i] == (uint64_t)(((u128)ca[i] + GL_P - cb[i]) % GL_P), "gl_sub");
        }
        // ---- every shift amount, on the first 2048 pairs ----
        const uint32_t m = std::min<uint32_t>(n, 2048);
        HIP_CHECK(hipFree(dout));
        HIP_CHECK(hipMalloc(&dout, 2ull * NPOW * m * sizeof(uint64_t)));
        hipLaunchKernelGGL(k_pow2, dim3((m + 255) / 256), block, 0, 0, da, dca, dcb, dout, m);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        const auto pw = host_copy(dout, 2ull * NPOW * m);
        for (int s = 0; s < NPOW; ++s)
            for (uint32_t i = 0; i < m; ++i) {
                CHECK(pw[(uint64_t)s * m + i] == gl_ref_pow2(a[i], s), "gl_mul_pow2<%d>(%016llx)", s,
                      (unsigned long long)a[i]);
                const uint64_t d = (uint64_t)(((u128)ca[i] + GL_P - cb[i]) % GL_P);
                CHECK(pw[(uint64_t)(NPOW + s) * m + i] == gl_ref_pow2(d, s), "gl_sub_mul_pow2<%d>", s);
            }
        HIP_CHECK(hipFree(da));
        HIP_CHECK(hipFree(db));
        HIP_CHECK(hipFree(dca));
        HIP_CHECK(hipFree(dcb));
        HIP_CHECK(hipFree(dout));
    }

    // ---- 64-bit Montgomery field ----
    for (uint64_t p : std::vector<uint64_t>{65537ull, 4294967291ull, 9223372036854775837ull, GL_P,
                                            18446744073709551557ull}) {
        const uint64_t pinv = m64_neg_inverse(p);
        std::vector<uint64_t> xa(n), xb(n);
        for (uint32_t i = 0; i < n; ++i) {
            xa[i] = a[i] % p;
            xb[i] = b[i] % p;
        }
        uint64_t *dxa = dev_copy(xa), *dxb = dev_copy(xb), *db = dev_copy(b), *dout = nullptr;
        HIP_CHECK(hipMalloc(&dout, 3ull * n * sizeof(uint64_t)));
        hipLaunchKernelGGL(k_m64, grid, block, 0, 0, dxa, db, dxa, dxb, dout, n, p, pinv);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        const auto out = host_copy(dout, 3ull * n);
        // m64_mul(a, b) = a b 2^-64 mod p: check out * 2^64 == a b (mod p)
        for (uint32_t i = 0; i < n; ++i) {
            CHECK(out[i] < p && (uint64_t)((((u128)out[i] << 64) % p)) == (uint64_t)((u128)xa[i] * b[i] % p),
                  "m64_mul p %llu a %016llx b %016llx", (unsigned long long)p, (unsigned long long)xa[i],
                  (unsigned long long)b[i]);
            CHECK(out[n + i] == (uint64_t)(((u128)xa[i] + xb[i]) % p), "m64_add p %llu", (unsigned long long)p);
            CHECK(out[2 * n + i] == (uint64_t)(((u128)xa[i] + p - xb[i]) % p), "m64_sub p %llu",
                  (unsigned long long)p);
        }
        HIP_CHECK(hipFree(dxa));
        HIP_CHECK(hipFree(dxb));
        HIP_CHECK(hipFree(db));
        HIP_CHECK(hipFree(dout));
    }

    // ---- 32-bit field ----
    for (uint32_t p : {3u, 17u, 65537u, 2013265921u, 2147483647u}) {
        const uint32_t pinv = u32_neg_inverse(p);
        std::vector<uint32_t> x(n), wm(n), xa(n), xb(n);
        for (uint32_t i = 0; i < n; ++i) {
            x[i] = (uint32_t)a[i];  // any 32-bit word
            wm[i] = (uint32_t)(b[i] % p);
            xa[i] = (uint32_t)(a[i] % p);
            xb[i] = (uint32_t)(b[i] % p);
        }
        uint32_t *dx = dev_copy(x), *dw = dev_copy(wm), *dxa = dev_copy(xa), *dxb = dev_copy(xb), *dout = nullptr;
        HIP_CHECK(hipMalloc(&dout, 3ull * n * sizeof(uint32_t)));
        hipLaunchKernelGGL(k_u32, grid, block, 0, 0, dx, dw, dxa, dxb, dout, n, p, pinv);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        const auto out = host_copy(dout, 3ull * n);
        for (uint32_t i = 0; i < n; ++i) {
            // u32_mont(x, w) = x w 2^-32 mod p
            CHECK(out[i] < p && (uint64_t)((((u128)out[i] << 32) % p)) == (uint64_t)((u128)x[i] * wm[i] % p),
                  "u32_mont p %u x %08x w %08x", p, x[i], wm[i]);
            CHECK(out[n + i] == (uint32_t)(((uint64_t)xa[i] + xb[i]) % p), "u32_add p %u", p);
            CHECK(out[2 * n + i] == (uint32_t)(((uint64_t)xa[i] + p - xb[i]) % p), "u32_sub p %u", p);
        }
        HIP_CHECK(hipFree(dx));
        HIP_CHECK(hipFree(dw));
        HIP_CHECK(hipFree(dxa));
        HIP_CHECK(hipFree(dxb));
        HIP_CHECK(hipFree(dout));
    }
    // ---- 32-bit field, lazy form (p < 2^30) ----
    for (uint32_t p : {3u, 17u, 7340033u, 998244353u, 1004535809u, 1073741789u}) {
        const uint32_t pinv = u32_neg_inverse(p), p2 = 2 * p;
        const uint32_t edge[] = {0, 1, p - 1, p, p + 1, p2 - 2, p2 - 1};
        std::vector<uint32_t> la(n), lb(n), wm(n);
        for (uint32_t i = 0; i < n; ++i) {
            la[i] = i < 49 ? edge[i % 7] : (uint32_t)(a[i] % p2);
            lb[i] = i < 49 ? edge[i / 7] : (uint32_t)(b[i] % p2);
            wm[i] = (uint32_t)((a[i] ^ b[i]) % p);
        }
        uint32_t *dla = dev_copy(la), *dlb = dev_copy(lb), *dw = dev_copy(wm), *dout = nullptr;
        HIP_CHECK(hipMalloc(&dout, 5ull * n * sizeof(uint32_t)));
        hipLaunchKernelGGL(k_u32l, grid, block, 0, 0, dla, dlb, dw, dout, n, p, pinv);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        const auto out = host_copy(dout, 5ull * n);
        const uint64_t r32 = (1ull << 32) % p;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t x = la[i], y = lb[i];
            CHECK(out[i] < p2 && out[i] % p == (x + y) % p, "u32l_add p %u", p);
            CHECK(out[n + i] < p2 && out[n + i] % p == (x + p2 - y) % p, "u32l_sub p %u", p);
            // r 2^32 == value (mod p), r < 2p
            CHECK(out[2 * n + i] < p2 && (uint64_t)((u128)out[2 * n + i] * r32 % p) ==
                                             (uint64_t)((u128)((x + p2 - y) % p) * wm[i] % p),
                  "u32l_sub_mont p %u", p);
            CHECK(out[3 * n + i] < p2 && (uint64_t)((u128)out[3 * n + i] * r32 % p) == (uint64_t)((u128)x * y % p),
                  "u32l_mont p %u", p);
            CHECK(out[4 * n + i] == x % p, "u32l_canon p %u", p);
        }
        HIP_CHECK(hipFree(dla));
        HIP_CHECK(hipFree(dlb));
        HIP_CHECK(hipFree(dw));
        HIP_CHECK(hipFree(dout));
    }
    std::printf("device field test: %ld cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
