// gb_big.hpp -- device side of the arbitrary-size single-n search
// (strong probable-prime tests of n - p, multi-word Montgomery arithmetic).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gbk {

constexpr uint32_t BIG_BS = 1024;      // threads per test, at most
constexpr uint32_t BIG_LMAX = 2304;    // 32-bit words of n: 73,728 bits, 22,195 decimal digits

struct BigTask {
    uint32_t p;      // q = n - p
    uint32_t base;   // strong-test base
};

// LDS bytes of one test workgroup for an n of L words, and its threads.
size_t big_lds_bytes(uint32_t L);
uint32_t big_threads(uint32_t L);

// pass[i] = 1 when n - tasks[i].p is a strong probable prime to base tasks[i].base
// (n: L >= 2 words, little-endian, n - p >= 2^32 so q keeps at least two words).
hipError_t launch_big_mr(hipStream_t s, const uint32_t* n, uint32_t L, const BigTask* tasks, uint32_t ntasks,
                         uint32_t* pass);

// bad: bit (p - 1)/2 set for every odd p <= p_max with some ell[i] dividing n - p
// (ell ascending, the first nsmall of them below 2^13; r: nell words of scratch;
// bad: zeroed, p_max / 64 + 1 words).
hipError_t launch_big_filter(hipStream_t s, const uint32_t* n, uint32_t L, const uint32_t* ell, uint32_t nell,
                             uint32_t nsmall, uint32_t* r, uint64_t p_max, uint32_t* bad);

}  // namespace gbk
