// test_bitset.cpp -- known-answer check of goldbach::build_prime_bitset:
// the primes <= 30 and pi(10^k) for k = 3..8 (the values the reference asserts
// in src/test_bitset.cpp:17-50), plus the tail-bit and word-count layout.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "prime_bitset.hpp"

using goldbach::PrimeBitset;

int main(int argc, char** argv) {
    if (argc == 3 && std::strcmp(argv[1], "--dump") == 0) {
        // words of build_prime_bitset(limit), one hex word per line
        auto bs = goldbach::build_prime_bitset(std::stoull(argv[2]));
        for (uint64_t i = 0; i < bs.word_count(); ++i) std::printf("%016llx\n", (unsigned long long)bs.data()[i]);
        return 0;
    }
    int bad = 0;
    {
        auto bs = goldbach::build_prime_bitset(30);
        std::vector<unsigned long long> got;
        for (unsigned long long n = 0; n <= 30; ++n)
            if (bs.is_prime(n)) got.push_back(n);
        const std::vector<unsigned long long> want = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29};
        const bool ok = got == want;
        std::printf("primes <= 30: %s\n", ok ? "PASS" : "FAIL");
        bad += !ok;
    }
    const unsigned long long pi[9] = {0, 4, 25, 168, 1229, 9592, 78498, 664579, 5761455};
    unsigned long long n = 100;
    for (int k = 3; k <= 8; ++k) {
        n *= 10;
        auto bs = goldbach::build_prime_bitset(n);
        unsigned long long c = bs.is_prime(2) ? 1 : 0;
        for (unsigned long long q = 3; q <= n; q += 2) c += bs.is_prime(q) ? 1 : 0;
        const bool ok = c == pi[k];
        std::printf("pi(%llu) = %llu expected %llu  %s\n", n, c, pi[k], ok ? "PASS" : "FAIL");
        bad += !ok;
    }
    {
        // limit 100: 49 odd bits, one word, bits 49..63 stay set
        auto bs = goldbach::build_prime_bitset(100);
        const bool ok = bs.word_count() == 1 && (bs.data()[0] >> 49) == 0x7FFFull &&
                        PrimeBitset::num_bits(100) == 49 && PrimeBitset::num_bits(2) == 0;
        std::printf("layout: %s\n", ok ? "PASS" : "FAIL");
        bad += !ok;
    }
    std::printf("\n%s\n", bad ? "SOME TESTS FAILED." : "All tests passed.");
    return bad ? 1 : 0;
}
