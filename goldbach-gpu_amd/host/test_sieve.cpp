// test_sieve.cpp -- known-answer check of simple_sieve / segmented_sieve:
// pi(10^k) for k = 1..8 (the values the reference asserts in src/test_sieve.cpp:31-40)
// plus a split-window consistency check.
#include <cstdio>
#include <cstring>
#include <string>

#include "segmented_sieve.hpp"

int main(int argc, char** argv) {
    if (argc == 4 && std::strcmp(argv[1], "--dump") == 0) {
        // segmented_sieve(low, high) as a string of 0/1
        for (char x : segmented_sieve(std::stoull(argv[2]), std::stoull(argv[3]))) std::putchar(x ? '1' : '0');
        std::putchar('\n');
        return 0;
    }
    const unsigned long long pi[9] = {0, 4, 25, 168, 1229, 9592, 78498, 664579, 5761455};
    int bad = 0;
    unsigned long long n = 1;
    for (int k = 1; k <= 8; ++k) {
        n *= 10;
        unsigned long long c = 0;
        for (char x : segmented_sieve(0, n)) c += x ? 1 : 0;
        const unsigned long long s = simple_sieve(n).size();
        const bool ok = c == pi[k] && s == pi[k];
        std::printf("pi(10^%d): segmented %llu simple %llu expected %llu  %s\n", k, c, s, pi[k], ok ? "PASS" : "FAIL");
        bad += !ok;
    }
    // [0, 10^6] split at an odd boundary must agree with the whole
    auto whole = segmented_sieve(0, 1000000);
    auto a = segmented_sieve(0, 499999), b = segmented_sieve(500000, 1000000);
    bool same = a.size() + b.size() == whole.size();
    for (size_t i = 0; same && i < a.size(); ++i) same = a[i] == whole[i];
    for (size_t i = 0; same && i < b.size(); ++i) same = b[i] == whole[a.size() + i];
    std::printf("split window: %s\n", same ? "PASS" : "FAIL");
    bad += !same;
    std::printf("\n%s\n", bad ? "SOME TESTS FAILED." : "All tests passed.");
    return bad ? 1 : 0;
}
