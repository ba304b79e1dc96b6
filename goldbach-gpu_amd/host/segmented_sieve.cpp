// segmented_sieve.cpp -- simple_sieve / segmented_sieve (CPU), drop-in for the
// reference's helpers (src/segmented_sieve.cpp:13-50): same signatures, same
// outputs (1 iff the integer is prime; 0 and 1 are not; a prime p inside
// [low, high] stays 1).
#include "segmented_sieve.hpp"

#include <algorithm>
#include <cmath>

namespace {
uint64_t floor_sqrt(uint64_t n) {
    uint64_t r = (uint64_t)std::sqrt((long double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) ++r;
    return r;
}
}  // namespace

std::vector<uint64_t> simple_sieve(uint64_t limit) {
    std::vector<uint64_t> primes;
    if (limit < 2) return primes;
    std::vector<uint8_t> composite(limit + 1, 0);
    for (uint64_t i = 2; i <= limit; ++i) {
        if (composite[i]) continue;
        primes.push_back(i);
        if (i <= limit / i)
            for (uint64_t j = i * i; j <= limit; j += i) composite[j] = 1;
    }
    return primes;
}

std::vector<char> segmented_sieve(uint64_t low, uint64_t high) {
    std::vector<char> is_prime(high - low + 1, 1);
    for (uint64_t v = low; v <= std::min<uint64_t>(high, 1); ++v) is_prime[v - low] = 0;
    for (uint64_t p : simple_sieve(floor_sqrt(high))) {
        uint64_t m = std::max(p * p, (low + p - 1) / p * p);
        for (uint64_t j = m; j <= high; j += p) {
            is_prime[j - low] = 0;
            if (high - j < p) break;
        }
    }
    return is_prime;
}
