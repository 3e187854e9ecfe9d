// dropin_host_test.cpp -- the host-only parts of the C++ drop-in
// (include/stntt/ntt_kernel.hpp): the reference's src/ntt/ntt_util.h helpers
// at global scope and cuda::ntt::NttArgs' launch geometry.  No GPU call is
// made (the default context is never touched), so this runs without a
// device.  Prints one "name args... = result" line per case;
// tests/test_dropin_host.py checks them against the reference's own
// functions (oracle/_ref) where those are built, and against exact integer
// arithmetic.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/stntt/ntt_kernel.hpp"

int main() {
    // elemWiseMul on signed ints, |a b| < 2^31 so the reference's int
    // product does not wrap
    uint64_t state = 12345;
    auto next = [&]() {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return (int64_t)(state >> 33);
    };
    for (int mod : {17, 12289, 40961, 65537, 7340033}) {
        std::vector<int> a(64), b(64);
        for (int i = 0; i < 64; ++i) {
            a[i] = (int)(next() % 80001) - 40000;
            b[i] = (int)(next() % 80001) - 40000;
        }
        std::vector<int> r = a;
        elemWiseMul(std::span<int>(r), std::span<int>(b), (size_t)mod);
        for (int i = 0; i < 64; ++i) std::printf("elemWiseMul %d %d %d = %d\n", a[i], b[i], mod, r[i]);
    }
    // modulo, pow (negative exponent = Fermat inverse), sqrt, isPrime
    for (int x : {-40, -17, -1, 0, 5, 16, 33})
        for (int m : {3, 16, 17}) std::printf("modulo %d %d = %d\n", x, m, modulo(x, m));
    for (long e : {-3L, -1L, 0L, 1L, 5L, 40L})
        for (long b : {-5L, 2L, 3L, 10L}) std::printf("pow %ld %ld %d = %ld\n", b, e, 7681, pow(b, e, 7681L));
    for (size_t v : {0ul, 1ul, 2ul, 15ul, 16ul, 17ul, 1000000ul, 18446744073709551615ul})
        std::printf("sqrt %zu = %zu\n", v, sqrt(v));
    for (size_t v : {0ul, 1ul, 2ul, 4ul, 7681ul, 12289ul, 65535ul, 65537ul})
        std::printf("isPrime %zu = %d\n", v, (int)isPrime(v));
    for (size_t v : {12ul, 7680ul, 65536ul, 2013265920ul}) {
        std::printf("uniquePrimeFactors %zu =", v);
        for (size_t f : uniquePrimeFactors(v)) std::printf(" %zu", f);
        std::printf("\n");
    }
    std::printf("findGenerator %d %d = %zu\n", 7680, 7681, findGenerator(7680, 7681));
    std::printf("findPrimitiveRoot %d %d %d = %zu\n", 512, 7680, 7681, findPrimitiveRoot(512, 7680, 7681));
    std::printf("isPrimitiveRoot %d %d %d = %d\n", 7146, 512, 7681, (int)isPrimitiveRoot(7146, 512, 7681));
    std::printf("findModulus %d %d = %zu\n", 1024, 11, findModulus(1024, 11));
    for (unsigned L : {1u, 9u, 20u}) {
        auto [root, mod] = findParams(1u << L, 11);
        std::printf("findParams %u %d = %zu %zu\n", 1u << L, 11, root, mod);
    }
    // bit reversal
    std::printf("bitReverse %d %d = %zu\n", 6, 3, bitReverse(6, 3));
    std::vector<int> v(16);
    for (int i = 0; i < 16; ++i) v[i] = i;
    bitReverseVector(v);
    std::printf("bitReverseVector 16 =");
    for (int x : v) std::printf(" %d", x);
    std::printf("\n");
    // the reference's launch geometry, as callers read it
    for (unsigned n : {4u, 64u, 1024u, 1u << 20})
        for (unsigned batches : {1u, 10u})
            for (unsigned bs : {64u, 1024u}) {
                cuda::ntt::NttArgs a(n, batches, 3, 17, 2, bs);
                std::printf("NttArgs %u %u %u = %u %u %u %u %u %u %u %d\n", n, batches, bs, a.dimGrid.x, a.dimGrid.y,
                            a.dimGrid.z, a.dimBlock.x, a.dimBlock.y, a.dimBlock.z, a.sharedMem, (int)a.isGlobal);
            }
    return 0;
}
