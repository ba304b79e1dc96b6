// cpu_goldbach_main.cpp -- the `cpu_goldbach` command line, drop-in for the
// reference's sequential CPU oracle (src/cpu_goldbach.cpp:45-130): same
// argument checks, messages, summary lines and exit codes, on the product's
// own segmented_sieve (libgoldbach.so).
//
// The reference tests p = 2, 3, 4, ... up to n/2 against a byte table for
// each even n in turn, on one thread, and stops at the first n without a
// partition (cpu_goldbach.cpp:14-24,98-111).  Here the n are cut into blocks
// scanned in parallel (OpenMP); each n tries only the primes p, ascending,
// so the first hit is the same minimal p and an n fails exactly when the
// reference's loop fails.  Blocks past a failure found so far are skipped,
// and the smallest failing n is the one reported, so "checked" and the FAIL
// line equal the sequential loop's.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <iostream>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "segmented_sieve.hpp"

namespace {

void print_usage(const char* prog) {
    std::cout << "Goldbach Sequential CPU Oracle\n";
    std::cout << "Usage: " << prog << " <LIMIT>\n";
    std::cout << "  LIMIT       Max even integer to check (e.g., 100000000)\n";
    std::cout << "  -h, --help  Show this help message\n";
}

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

constexpr uint64_t NBLOCK = 1ull << 16;   // even n per scheduling block

// Does n (even, >= 4) have a partition p + q, p prime <= n/2, q prime?  The odd primes come
// from `odd` (ascending); past its end the table itself is scanned.
bool has_partition(uint64_t n, const std::vector<char>& is_prime, const std::vector<uint32_t>& odd) {
    if (n == 4) return true;                       // 2 + 2; no other even n has p = 2
    const uint64_t half = n / 2;
    for (uint32_t p : odd) {
        if (p > half) return false;
        if (is_prime[n - p]) return true;
    }
    for (uint64_t p = (odd.empty() ? 3 : (uint64_t)odd.back() + 2); p <= half; p += 2)
        if (is_prime[p] && is_prime[n - p]) return true;
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2 || std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help") {
        print_usage(argv[0]);
        return 0;
    }
    uint64_t LIMIT;
    try {
        LIMIT = std::stoull(argv[1]);
    } catch (...) {
        std::cerr << "Error: Invalid numeric argument.\n";
        return 1;
    }
    if (LIMIT < 4) {
        std::cerr << "Error: LIMIT must be >= 4.\n";
        return 1;
    }
    if (LIMIT % 2 != 0) LIMIT--;

    std::cout << "Building prime table up to " << LIMIT << "...\n";
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<char> is_prime;
    try {
        is_prime = segmented_sieve(0, LIMIT);
    } catch (const std::bad_alloc&) {
        std::cerr << "\n[!] ERROR: System RAM exhausted.\n";
        std::cerr << "[!] A limit of " << LIMIT << " requires ~" << LIMIT / (1024 * 1024 * 1024)
                  << " GB of RAM for the CPU byte-array.\n";
        std::cerr << "[!] SOLUTION: Use 'goldbach' (GPU, bitsets and segments) for limits this large.\n";
        return 1;
    } catch (const std::length_error&) {
        std::cerr << "\n[!] ERROR: System RAM exhausted.\n";
        std::cerr << "[!] A limit of " << LIMIT << " exceeds the largest byte-array this system can index.\n";
        std::cerr << "[!] SOLUTION: Use 'goldbach' (GPU, bitsets and segments) for limits this large.\n";
        return 1;
    }
    const double sieve_ms = ms_since(t0);
    std::cout << "Sieve done in " << sieve_ms << " ms\n";

    std::cout << "Checking Goldbach for all even n in [4, " << LIMIT << "]...\n";
    const auto t2 = std::chrono::steady_clock::now();
    // odd primes up to the largest p any minimal partition below 2^64 needs, and then some
    std::vector<uint32_t> odd;
    const uint64_t plist = std::min<uint64_t>(LIMIT / 2, 1u << 20);
    for (uint64_t p = 3; p <= plist; p += 2)
        if (is_prime[p]) odd.push_back((uint32_t)p);

    const uint64_t evens = (LIMIT - 4) / 2 + 1;
    const int64_t nblk = (int64_t)((evens + NBLOCK - 1) / NBLOCK);
    std::atomic<uint64_t> first_fail{UINT64_MAX};
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nblk; ++b) {
        const uint64_t n_lo = 4 + 2 * (uint64_t)b * NBLOCK;
        if (n_lo > first_fail.load(std::memory_order_relaxed)) continue;
        const uint64_t n_hi = std::min(LIMIT, n_lo + 2 * (NBLOCK - 1));
        for (uint64_t n = n_lo; n <= n_hi; n += 2) {
            if (has_partition(n, is_prime, odd)) continue;
            uint64_t cur = first_fail.load();
            while (n < cur && !first_fail.compare_exchange_weak(cur, n)) {
            }
            break;
        }
    }
    uint64_t failures = 0, checked = evens;
    if (first_fail.load() != UINT64_MAX) {
        failures = 1;
        checked = (first_fail.load() - 4) / 2 + 1;
        std::cout << "FAIL: Goldbach fails at n = " << first_fail.load() << "\n";
    }
    const double check_ms = ms_since(t2);

    std::cout << "\n--- Summary ---\n";
    std::cout << "Even numbers checked : " << checked << "\n";
    std::cout << "Failures found       : " << failures << "\n";
    std::cout << "Sieve time           : " << sieve_ms << " ms\n";
    std::cout << "Goldbach check time  : " << check_ms << " ms\n";
    std::cout << "Total time           : " << sieve_ms + check_ms << " ms\n";
    if (failures == 0) std::cout << "\nAll even numbers up to " << LIMIT << " satisfy Goldbach. ✓\n";
    return failures == 0 ? 0 : 1;
}
