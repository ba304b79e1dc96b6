// prime_bitset.cpp -- goldbach::build_prime_bitset (CPU, OpenMP).
// Same result and layout as the reference's builder (src/prime_bitset.cpp:21-97);
// work is split into slices of whole 64-bit words, so no two threads ever
// update the same word (the reference's slices are not word-aligned).
#include "prime_bitset.hpp"

#include <algorithm>
#include <cmath>

namespace goldbach {

namespace {
uint64_t floor_sqrt(uint64_t n) {
    uint64_t r = (uint64_t)std::sqrt((long double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) ++r;
    return r;
}
}  // namespace

PrimeBitset build_prime_bitset(uint64_t limit) {
    PrimeBitset bs(limit);
    if (limit < 9) return bs;   // 3, 5, 7 are prime: nothing to clear
    const uint64_t root = floor_sqrt(limit);
    std::vector<uint8_t> composite(root + 1, 0);
    std::vector<uint64_t> base;
    for (uint64_t i = 3; i <= root; i += 2) {
        if (composite[i]) continue;
        base.push_back(i);
        for (uint64_t j = i * i; j <= root; j += 2 * i) composite[j] = 1;
    }
    const uint64_t nbits = PrimeBitset::num_bits(limit);
    const uint64_t nwords = bs.word_count();
    uint64_t* words = bs.data();
    const uint64_t slice = 1ull << 15;   // words per task (256 KiB)
    const int64_t nslices = (int64_t)((nwords + slice - 1) / slice);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < nslices; ++s) {
        const uint64_t w0 = (uint64_t)s * slice;
        const uint64_t w1 = std::min(nwords, w0 + slice);
        const uint64_t b1 = std::min(nbits, w1 * 64);
        const uint64_t lo = 2 * (w0 * 64) + 3, hi = 2 * (b1 - 1) + 3;
        for (uint64_t p : base) {
            const uint64_t pp = p * p;
            if (pp > hi) break;
            uint64_t m = pp;
            if (m < lo) {
                m = (lo + p - 1) / p * p;
                if ((m & 1) == 0) m += p;
            }
            for (uint64_t j = m; j <= hi; j += 2 * p) {
                const uint64_t b = (j - 3) >> 1;
                words[b >> 6] &= ~(1ull << (b & 63));
                if (hi - j < 2 * p) break;
            }
        }
    }
    return bs;
}

}  // namespace goldbach
