/*
 * oracle.h -- CPU restatement of the reference's segmented double-sieve
 * Goldbach range verifier (isaac-6/goldbach-gpu).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker
 * (or, for the baseline, as the timed CPU restatement).  The product under
 * goldbach-gpu_amd/ never links, loads or calls it.
 *
 * Every function is exact integer arithmetic: primality comes from a sieve
 * or from deterministic Miller-Rabin with the 12 bases 2..37 (valid for all
 * n < 3.3e24, Sorenson & Webster 2015), so no result is ever approximated.
 *
 * Semantics restated (reference file:line):
 *   orc_simple_sieve        src/segmented_sieve.cpp:13-26
 *   orc_segmented_sieve     src/segmented_sieve.cpp:29-50
 *   orc_prime_bitset        include/prime_bitset.hpp:29-101 + src/prime_bitset.cpp:21-97
 *                           (race-free; tail bits beyond num_bits stay 1, hpp:38)
 *   orc_window_bits         src/goldbach.cu:198-259 (bit i <-> q_low + 2i, tail bits 1)
 *   orc_is_prime            src/goldbach.cu:333-360 (also :119-158, single_check.cu:30-88)
 *   orc_min_partition       src/cpu_goldbach.cpp:14-24 scan order (first p ascending)
 *   orc_verify_segment      src/goldbach.cu:288-311 (Phase 1), :362-373 (Phase 2),
 *                           :420-485 (per-segment loop), :160-179 (q resolution)
 *   orc_verify_range        src/goldbach.cu:420-433, 689, 786-801 (segment walk + summary)
 *   orc_cpu_goldbach        src/cpu_goldbach.cpp:45-130
 *   orc_count_partitions    src/cpu_goldbach.cpp:29-36 (goldbach_count)
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   pi(10^k) k=1..8 and primes <= 30  -- src/test_sieve.cpp:31-40, src/test_bitset.cpp:17-50
 *   byte/word equality with the reference's own segmented_sieve / build_prime_bitset
 *     compiled from the reference tree into oracle/_ref (where that tree is present)
 *   cpu_goldbach binary from oracle/_ref on the same LIMIT
 *   minimal-p record table to 1e8 (OEIS A025018/A025019) and literature p_min values
 *   published partitions RESULTS.md:114-121 (both parts prime)
 *   brute-force trial division on small inputs; MR vs sieve on overlapping ranges
 */
#ifndef GB_ORACLE_H
#define GB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_stats {
    uint64_t evens;        /* even n examined */
    uint64_t phase2;       /* n with no prime p <= p_small giving n-p prime */
    uint64_t failures;     /* n with no prime p <= max(p_small, phase2_limit) */
    uint64_t first_fail_n; /* smallest failing n, 0 if none */
    uint64_t sum_pmin;     /* sum of p_min(n) over n with a partition found */
    uint64_t max_pmin;     /* largest p_min(n) */
    uint64_t argmax_n;     /* smallest n with p_min(n) == max_pmin */
} orc_stats;

uint64_t orc_isqrt(uint64_t n);

/* primes <= limit; returns the count, writes min(count, cap) primes to out (out may be NULL). */
size_t orc_simple_sieve(uint64_t limit, uint64_t *out, size_t cap);

/* out[i] = 1 iff low+i is prime, for i in [0, high-low].  Returns 0, or -1 on bad args/alloc. */
int orc_segmented_sieve(uint64_t low, uint64_t high, uint8_t *out);

/* PrimeBitset(limit): number of 64-bit words, and the words after build_prime_bitset. */
uint64_t orc_bitset_word_count(uint64_t limit);
void orc_prime_bitset(uint64_t limit, uint64_t *words);

/* Window bitset over odd q in [q_low, q_high] (both odd, 3 <= q_low <= q_high). */
uint64_t orc_window_word_count(uint64_t q_low, uint64_t q_high);
int orc_window_bits(uint64_t q_low, uint64_t q_high, uint64_t *words);

int orc_is_prime(uint64_t n);
/* n passes the strong (Miller-Rabin) test to each of the first k of the bases 2..37 */
int orc_strong_first(uint64_t n, int k);
/* smallest odd composite in [9, limit] passing the first k bases (0 if none): brute force */
uint64_t orc_first_strong_pseudoprime(int k, uint64_t limit);

/* Smallest prime p <= min(bound, n/2) with n-p prime; 0 if none (or n odd / n < 4). */
uint64_t orc_min_partition(uint64_t n, uint64_t bound);
void orc_min_partition_batch(const uint64_t *n, size_t count, uint64_t bound, uint64_t *p_out, int threads);

/* Goldbach partition counts: out[i] = #{p prime, p <= n/2, n - p prime} for n = start + 2i,
 * i < count (cpu_goldbach.cpp:29-36 goldbach_count: p from 2 to n/2 over a byte table).
 * One sieve of [0, n_max] then a scan per n (OpenMP over n).  Returns 0, or -1 when start is
 * odd or below 4 or the table cannot be allocated. */
int orc_count_partitions(uint64_t start, uint64_t count, uint64_t *out, int threads);

/*
 * One segment of `count` consecutive even n starting at even A >= 4.
 * verified_bits (nullable): ceil(count/64) words, bit j of word w <-> n = A + 2(64w+j),
 *   set iff p_min(n) <= p_small (the reference's Phase-1 verified flag); tail bits 0.
 * pmin_out (nullable): count entries, p_min(n) searched up to max(p_small, phase2_limit),
 *   0 when no partition was found below that bound.
 */
int orc_verify_segment(uint64_t A, uint64_t count, uint64_t p_small, uint64_t phase2_limit,
                       uint64_t *verified_bits, uint32_t *pmin_out, orc_stats *st);

/* All even n in [start, limit] walked in segments of seg_evens evens (OpenMP over segments). */
int orc_verify_range(uint64_t start, uint64_t limit, uint64_t seg_evens, uint64_t p_small,
                     uint64_t phase2_limit, int threads, orc_stats *st);

/* cpu_goldbach LIMIT: checked count, failures (stops at the first), first failing n. */
int orc_cpu_goldbach(uint64_t limit, uint64_t *checked, uint64_t *failures, uint64_t *first_fail);

#ifdef __cplusplus
}
#endif
#endif
