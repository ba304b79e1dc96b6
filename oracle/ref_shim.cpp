// ref_shim.cpp -- extern "C" entry points over the reference's own host
// library functions, so tests can call them through ctypes.  Built only into
// oracle/_ref/ (never shipped): the reference sources are compiled where they
// lie in the reference tree by oracle/Makefile; nothing of them is copied.
//
// Wrapped (reference file:line):
//   simple_sieve(uint64_t)                 src/segmented_sieve.cpp:13-26
//   segmented_sieve(uint64_t, uint64_t)    src/segmented_sieve.cpp:29-50
//   goldbach::build_prime_bitset(uint64_t) src/prime_bitset.cpp:21-97
//   goldbach_count(uint64_t, is_prime)     src/cpu_goldbach.cpp:29-36 (that file is
//                                          compiled with its main renamed to cpu_goldbach_main)
// build_prime_bitset races on shared boundary words when run with more than
// two OpenMP threads, so the caller picks the thread count (1 for parity).
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>
#include "prime_bitset.hpp"

std::vector<uint64_t> simple_sieve(uint64_t limit);
std::vector<char> segmented_sieve(uint64_t low, uint64_t high);
uint64_t goldbach_count(uint64_t n, const std::vector<char>& is_prime);

extern "C" {

size_t ref_simple_sieve(uint64_t limit, uint64_t* out, size_t cap) {
    std::vector<uint64_t> v = simple_sieve(limit);
    size_t n = v.size() < cap ? v.size() : cap;
    if (out && n) std::memcpy(out, v.data(), n * sizeof(uint64_t));
    return v.size();
}

int ref_segmented_sieve(uint64_t low, uint64_t high, uint8_t* out) {
    std::vector<char> v = segmented_sieve(low, high);
    std::memcpy(out, v.data(), v.size());
    return 0;
}

uint64_t ref_prime_bitset(uint64_t limit, uint64_t* out, uint64_t cap, int threads) {
    omp_set_num_threads(threads > 0 ? threads : 1);
    goldbach::PrimeBitset b = goldbach::build_prime_bitset(limit);
    uint64_t n = b.word_count() < cap ? b.word_count() : cap;
    if (out && n) std::memcpy(out, b.data(), n * sizeof(uint64_t));
    return b.word_count();
}

int ref_bitset_is_prime(uint64_t limit, const uint64_t* queries, uint8_t* out, uint64_t count) {
    omp_set_num_threads(1);
    goldbach::PrimeBitset b = goldbach::build_prime_bitset(limit);
    for (uint64_t i = 0; i < count; ++i) out[i] = b.is_prime(queries[i]) ? 1 : 0;
    return 0;
}

// out[i] = goldbach_count(start + 2i) over the reference's own byte sieve of [0, n_max].
int ref_goldbach_count(uint64_t start, uint64_t count, uint64_t* out) {
    if (count == 0) return 0;
    const uint64_t n_max = start + 2 * (count - 1);
    std::vector<char> is_prime = segmented_sieve(0, n_max);
    for (uint64_t i = 0; i < count; ++i) out[i] = goldbach_count(start + 2 * i, is_prime);
    return 0;
}

}  // extern "C"
