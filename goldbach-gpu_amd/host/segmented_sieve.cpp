// segmented_sieve.cpp -- simple_sieve / segmented_sieve (CPU), drop-in for the
// reference's helpers (src/segmented_sieve.cpp:13-50): same signatures, same
// outputs (1 iff the integer is prime; 0 and 1 are not; a prime p inside
// [low, high] stays 1).  The reference marks the whole byte array prime by
// prime; here [low, high] is cut into blocks that stay in a core's cache,
// and the blocks are sieved in parallel (OpenMP), each by every base prime.
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

constexpr uint64_t BLOCK = 1ull << 18;   // bytes of the result one thread sieves at a time
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
    std::vector<char> is_prime(high - low + 1, 1);   // std::bad_alloc / std::length_error propagate
    for (uint64_t v = low; v <= std::min<uint64_t>(high, 1); ++v) is_prime[v - low] = 0;
    const std::vector<uint64_t> base = simple_sieve(floor_sqrt(high));
    const uint64_t len = high - low + 1;
    const int64_t nblk = (int64_t)((len + BLOCK - 1) / BLOCK);
    char* out = is_prime.data();
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = 0; b < nblk; ++b) {
        const uint64_t lo = low + (uint64_t)b * BLOCK;
        const uint64_t hi = std::min(high, lo + (BLOCK - 1));
        for (uint64_t p : base) {
            if (p > hi / p) break;                  // p^2 > hi: no composite of this block has p as least factor
            const uint64_t r = lo % p;
            if (r && p - r > hi - lo) continue;     // no multiple of p in [lo, hi]
            const uint64_t m = std::max(p * p, r ? lo + (p - r) : lo);
            for (uint64_t j = m;; j += p) {
                out[j - low] = 0;
                if (hi - j < p) break;
            }
        }
    }
    return is_prime;
}
