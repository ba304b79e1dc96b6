// prime_bitset.hpp -- odd-only prime bitset, drop-in for the reference's
// goldbach::PrimeBitset / goldbach::build_prime_bitset
// (reference: include/prime_bitset.hpp:29-107, src/prime_bitset.cpp:21-97).
//
// Layout (unchanged, other code reads data() directly):
//   bit i of the word array  <->  odd n = 2i + 3,  word = i / 64, bit = i % 64 (LSB first)
//   num_bits(limit) = (limit - 3)/2 + 1 for limit >= 3, else 0
//   bits past num_bits in the last word are 1 (the constructor fills ~0 and the
//   builder only clears composites <= limit)
//
// build_prime_bitset here gives every worker thread whole 64-bit words, so no
// word is ever written by two threads.  The same words built on the GPU come
// from the C ABI: gb_sieve_window(ctx, 3, q_high, words) with q_high the
// largest odd number <= limit (include/goldbach_gpu.h).
#pragma once

#include <cassert>
#include <cstdint>
#include <vector>

namespace goldbach {

class PrimeBitset {
public:
    explicit PrimeBitset(uint64_t limit) : limit_(limit), words_((num_bits(limit) + 63) / 64, ~0ULL) {}

    // 2 is prime, other evens are not, n > limit reports false.
    [[nodiscard]] bool is_prime(uint64_t n) const {
        if (n == 2) return true;
        if (n < 3 || (n & 1) == 0 || n > limit_) return false;
        return test(n);
    }

    // Bit access for odd n >= 3 (asserted, as in the reference).
    void set(uint64_t n) {
        assert((n & 1) == 1 && n >= 3 && "set: n must be odd and >= 3");
        words_[index(n) >> 6] |= bit(n);
    }
    void clear(uint64_t n) {
        assert((n & 1) == 1 && n >= 3 && "clear: n must be odd and >= 3");
        words_[index(n) >> 6] &= ~bit(n);
    }
    bool test(uint64_t n) const {
        assert((n & 1) == 1 && n >= 3 && "test: n must be odd and >= 3");
        return (words_[index(n) >> 6] & bit(n)) != 0;
    }

    uint64_t limit() const { return limit_; }
    uint64_t word_count() const { return words_.size(); }
    const uint64_t* data() const { return words_.data(); }
    uint64_t* data() { return words_.data(); }
    uint64_t memory_bytes() const { return words_.size() * sizeof(uint64_t); }

    [[nodiscard]] static constexpr uint64_t num_bits(uint64_t limit) { return limit < 3 ? 0 : (limit - 3) / 2 + 1; }

private:
    static constexpr uint64_t index(uint64_t n) { return (n - 3) >> 1; }
    static constexpr uint64_t bit(uint64_t n) { return 1ULL << (index(n) & 63); }

    uint64_t limit_;
    std::vector<uint64_t> words_;
};

// Fully sieved bitset of the odd primes <= limit (CPU, OpenMP, race-free).
PrimeBitset build_prime_bitset(uint64_t limit);

}  // namespace goldbach
