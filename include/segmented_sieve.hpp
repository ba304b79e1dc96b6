// segmented_sieve.hpp -- the reference's plain sieve helpers, declared in a
// header (the reference declares them ad hoc in each consumer:
// src/cpu_goldbach.cpp:7, src/single_check.cu:15, src/test_sieve.cpp:6-7).
//
//   simple_sieve(limit)          every prime <= limit, ascending
//                                (reference src/segmented_sieve.cpp:13-26)
//   segmented_sieve(low, high)   one char per integer of [low, high]: 1 iff prime
//                                (reference src/segmented_sieve.cpp:29-50)
// Both return by value; std::bad_alloc propagates to the caller.
#pragma once

#include <cstdint>
#include <vector>

std::vector<uint64_t> simple_sieve(uint64_t limit);
std::vector<char> segmented_sieve(uint64_t low, uint64_t high);
