// field_test.cpp -- host-side check of the device field arithmetic
// (stntt_amd/csrc/field_gl.hpp, field_u32.hpp, field_m64.hpp) against 128-bit integer
// arithmetic on edge values and random values.  Runs on the CPU.
#include <cstdio>
#include <random>
#include <vector>

#include "../../stntt_amd/csrc/field_gl.hpp"
#include "../../stntt_amd/csrc/field_m64.hpp"
#include "../../stntt_amd/csrc/field_u32.hpp"

using namespace stntt;
typedef unsigned __int128 u128;

static int fails = 0;
#define CHECK(cond, ...) do { if (!(cond)) { if (fails < 20) { printf(__VA_ARGS__); printf("\n"); } ++fails; } } while (0)

static uint64_t ref_mul(uint64_t a, uint64_t b) { return (uint64_t)((u128)a * b % GL_P); }
static uint64_t ref_pow2(uint64_t x, int s) {
    u128 r = x % GL_P;
    for (int i = 0; i < s; ++i) r = (r * 2) % GL_P;
    return (uint64_t)r;
}

template <int S>
static void check_pow2(const std::vector<uint64_t> &vals) {
    for (uint64_t x : vals) {
        uint64_t got = gl_mul_pow2<S>(x), want = ref_pow2(x, S);
        CHECK(got == want, "mul_pow2<%d>(%llx) = %llx want %llx", S, (unsigned long long)x, (unsigned long long)got,
              (unsigned long long)want);
    }
}

template <int... S>
static void check_all_pow2(const std::vector<uint64_t> &vals, std::integer_sequence<int, S...>) {
    (check_pow2<S>(vals), ...);
}

int main() {
    std::vector<uint64_t> any = {0, 1, 2, 3, GL_EPS - 1, GL_EPS, GL_EPS + 1, 1ull << 32, (1ull << 32) + 1,
                                 GL_P - 2, GL_P - 1, GL_P, GL_P + 1, ~0ull, ~0ull - 1, 1ull << 63, (1ull << 63) - 1,
                                 0xFFFFFFFF00000000ull, 0x00000000FFFFFFFEull, 0x8000000080000000ull};
    std::mt19937_64 rng(12345);
    for (int i = 0; i < 3000; ++i) any.push_back(rng());
    for (int i = 0; i < 500; ++i) any.push_back(GL_P - 1 - (rng() & 0xFFFFF));
    for (int i = 0; i < 500; ++i) any.push_back(~0ull - (rng() & 0xFFFFFFFF));
    std::vector<uint64_t> canon;
    for (uint64_t x : any) canon.push_back(x % GL_P);

    for (uint64_t x : any) CHECK(gl_canon(x) == x % GL_P, "canon %llx", (unsigned long long)x);
    for (size_t i = 0; i < canon.size(); ++i)
        for (size_t j = 0; j < canon.size(); j += 37) {
            uint64_t a = canon[i], b = canon[j];
            CHECK(gl_add(a, b) == (uint64_t)(((u128)a + b) % GL_P), "add %llx %llx", (unsigned long long)a,
                  (unsigned long long)b);
            CHECK(gl_sub(a, b) == (uint64_t)(((u128)a + GL_P - b) % GL_P), "sub %llx %llx", (unsigned long long)a,
                  (unsigned long long)b);
        }
    for (size_t i = 0; i < any.size(); ++i)
        for (size_t j = 0; j < any.size(); j += 41) {
            uint64_t a = any[i], b = any[j];
            CHECK(gl_mul(a, b) == ref_mul(a, b), "mul %llx %llx", (unsigned long long)a, (unsigned long long)b);
            uint64_t lo = a, hi = b;
            uint64_t want = (uint64_t)((((u128)hi << 64) | lo) % GL_P);
            CHECK(gl_reduce128(lo, hi) == want, "reduce128 %llx %llx", (unsigned long long)lo, (unsigned long long)hi);
        }
    check_all_pow2(any, std::make_integer_sequence<int, 192>{});
    // sign folding: (a - b) * 2^S
    for (size_t i = 0; i < canon.size(); i += 7) {
        uint64_t a = canon[i], b = canon[(i * 13 + 5) % canon.size()];
        uint64_t d = (uint64_t)(((u128)a + GL_P - b) % GL_P);
        CHECK(gl_sub_mul_pow2<120>(a, b) == ref_pow2(d, 120), "submul120");
        CHECK(gl_sub_mul_pow2<36>(a, b) == ref_pow2(d, 36), "submul36");
        CHECK(gl_sub_mul_pow2<96>(a, b) == ref_pow2(d, 96), "submul96");
        // 64 <= s < 96 directly and after the sign fold (S - 96 in that range)
        CHECK(gl_sub_mul_pow2<64>(a, b) == ref_pow2(d, 64), "submul64");
        CHECK(gl_sub_mul_pow2<72>(a, b) == ref_pow2(d, 72), "submul72");
        CHECK(gl_sub_mul_pow2<95>(a, b) == ref_pow2(d, 95), "submul95");
        CHECK(gl_sub_mul_pow2<168>(a, b) == ref_pow2(d, 168), "submul168");
        CHECK(gl_sub_mul_pow2<180>(a, b) == ref_pow2(d, 180), "submul180");
        CHECK(gl_sub_mul_pow2<0>(a, b) == d, "submul0");
    }
    // 32-bit field: Montgomery and Barrett products, add, sub, input reduction,
    // for the findParams primes and the largest prime below 2^31
    for (uint32_t p : {11u, 12289u, 65537u, 7340033u, 2013265921u, 2147483629u}) {
        uint32_t bar = (uint32_t)((1ull << 32) / p), c32 = (uint32_t)((1ull << 32) % p);
        const uint32_t pinv = u32_neg_inverse(p);
        CHECK(p * pinv == 0xFFFFFFFFu, "u32 -p^-1 for %u", p);
        CHECK(u32_to_mont(1, p) == c32, "u32 Montgomery one for %u", p);
        std::vector<uint32_t> xs = {0, 1, p - 1, p / 2, 0xFFFFFFFFu, 0x80000000u, 0x7FFFFFFFu};
        for (int i = 0; i < 400; ++i) xs.push_back((uint32_t)rng());
        for (uint32_t x : xs) {
            uint32_t canon_x = x % p;
            CHECK(u32_reduce_input(x, p, bar, c32, false) == canon_x, "u32 reduce %u mod %u", x, p);
            int64_t sx = (int32_t)x;
            uint32_t want_s = (uint32_t)(((sx % (int64_t)p) + p) % p);
            CHECK(u32_reduce_input(x, p, bar, c32, true) == want_s, "u32 signed reduce %u mod %u", x, p);
            for (uint32_t y : xs) {
                uint32_t a = x % p, b = y % p;
                CHECK(u32_add(a, b, p) == (uint32_t)(((uint64_t)a + b) % p), "u32 add");
                CHECK(u32_sub(a, b, p) == (uint32_t)(((uint64_t)a + p - b) % p), "u32 sub");
                // any 32-bit data word times a Montgomery twiddle; twiddle times twiddle
                const uint32_t bm = u32_to_mont(b, p), am = u32_to_mont(a, p);
                CHECK(u32_mont(x, bm, p, pinv) == (uint32_t)((uint64_t)x * b % p), "u32 mont %u * %u mod %u", x, b, p);
                CHECK(u32_mont(am, bm, p, pinv) == u32_to_mont((uint64_t)a * b % p, p), "u32 mont twiddle product");
                CHECK(u32_mont(bm, 1, p, pinv) == b, "u32 mont to normal form");
                CHECK(u32_mul_any(a, b, p, ~0ull / p) == (uint32_t)((uint64_t)a * b % p), "u32 mul_any %u * %u mod %u",
                      a, b, p);
            }
        }
    }
    // 32-bit field, lazy form (p < 2^30): operands anywhere in [0, 2p),
    // results in [0, 2p) and congruent to the exact value
    for (uint32_t p : {3u, 11u, 65537u, 7340033u, 998244353u, 1004535809u, 1053818881u, 1073741789u}) {
        CHECK(u32_lazy_modulus(p), "lazy modulus %u", p);
        const uint32_t pinv = u32_neg_inverse(p), p2 = 2 * p;
        std::vector<uint32_t> ls = {0, 1, p - 1, p, p + 1, p2 - 2, p2 - 1, p / 2, p + p / 2};
        for (int i = 0; i < 160; ++i) ls.push_back((uint32_t)(rng() % p2));
        std::vector<uint32_t> ws = {0, 1, p - 1, p / 2};
        for (int i = 0; i < 12; ++i) ws.push_back((uint32_t)(rng() % p));
        const uint64_t r32 = (1ull << 32) % p;
        auto mont_ok = [&](uint32_t r, uint64_t prod) {  // r 2^32 == prod (mod p)
            return r < p2 && (uint64_t)(((u128)r * r32) % p) == (uint64_t)(prod % p);
        };
        for (uint32_t a : ls) {
            CHECK(u32l_canon(a, p) == a % p, "u32l canon %u mod %u", a, p);
            for (uint32_t b : ls) {
                const uint32_t s = u32l_add(a, b, p2), d = u32l_sub(a, b, p2);
                CHECK(s < p2 && s % p == (uint32_t)(((uint64_t)a + b) % p), "u32l add %u %u mod %u", a, b, p);
                CHECK(d < p2 && d % p == (uint32_t)(((uint64_t)a + p2 - b) % p), "u32l sub %u %u mod %u", a, b, p);
                // two lazy words (running twiddle products), and u32_mul_any on lazy words
                CHECK(mont_ok(u32l_mont(a, b, p, pinv), (u128)a * b % p), "u32l mont %u %u mod %u", a, b, p);
                CHECK(u32_mul_any(a, b, p, ~0ull / p) == (uint32_t)((uint64_t)a * b % p), "u32 mul_any lazy %u %u", a, b);
                for (uint32_t w : ws) {
                    const uint64_t diff = ((uint64_t)a + p2 - b) % p;
                    CHECK(mont_ok(u32l_sub_mont(a, b, w, p, pinv), (u128)diff * w), "u32l sub_mont %u %u %u mod %u",
                          a, b, w, p);
                }
            }
            // a biased difference (below 4p) times a table word
            for (uint32_t w : ws) {
                const uint32_t x = a + p2 - 1;
                CHECK(mont_ok(u32l_mont(x, w, p, pinv), (u128)x * w), "u32l mont 4p %u %u mod %u", x, w, p);
            }
        }
    }
    CHECK(!u32_lazy_modulus(1085276161u) && !u32_lazy_modulus(2147483647u), "lazy modulus bound");
    // 64-bit Montgomery field: primes below 2^32, between 2^32 and 2^63,
    // above 2^63, Goldilocks and the largest 64-bit prime
    for (uint64_t p : std::vector<uint64_t>{65537ull, 4294967291ull, 1099511627791ull, 0x3FFFFFFFFFFFFFFDull,
                                            9223372036854775837ull, GL_P, 18446744073709551557ull}) {
        uint64_t pinv = m64_neg_inverse(p);
        CHECK(p * pinv == (uint64_t)0 - 1, "m64 -p^-1 for %llu", (unsigned long long)p);
        uint64_t r1 = (uint64_t)(((u128)1 << 64) % p), r2 = (uint64_t)((u128)r1 * r1 % p);
        std::vector<uint64_t> xs = {0, 1, 2, p - 1, p - 2, p / 2, p / 2 + 1, 0xFFFFFFFFull, 1ull << 32};
        for (int i = 0; i < 300; ++i) xs.push_back(rng() % p);
        for (uint64_t &x : xs) x %= p;  // add and sub take canonical words
        std::vector<uint64_t> raw = {p, p + 1, ~0ull, ~0ull - 1, 1ull << 63};
        for (int i = 0; i < 100; ++i) raw.push_back(rng());
        auto mont = [&](uint64_t x) { return (uint64_t)(((u128)(x % p) << 64) % p); };
        for (uint64_t a : xs) {
            for (uint64_t b : xs) {
                CHECK(m64_add(a, b, p) == (uint64_t)(((u128)a + b) % p), "m64 add");
                CHECK(m64_sub(a, b, p) == (uint64_t)(((u128)a + p - b) % p), "m64 sub");
                // mont(a R, b) = a b: a product in the transform's convention
                CHECK(m64_mul(mont(a), b, p, pinv) == (uint64_t)((u128)a * b % p), "m64 mul %llu %llu mod %llu",
                      (unsigned long long)a, (unsigned long long)b, (unsigned long long)p);
            }
        }
        // any 64-bit word converts exactly: mont(x, R^2) = x R mod p; back with mont(., 1)
        for (uint64_t x : raw) {
            uint64_t xm = m64_mul(x, r2, p, pinv);
            CHECK(xm == mont(x), "m64 to-Montgomery of %llu mod %llu", (unsigned long long)x, (unsigned long long)p);
            CHECK(m64_mul(xm, 1, p, pinv) == x % p, "m64 from-Montgomery");
        }
    }
    // 2^39 is a primitive 64th root of unity (omega_64 for generator 7)
    CHECK(ref_pow2(1, 39 * 32) == GL_P - 1, "omega_64^32 != -1");
    printf("field test: %d failures\n", fails);
    return fails ? 1 : 0;
}
