/*
 * oracle.c -- CPU restatement of the reference's Goldbach range verifier.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): loaded by tests/, smoke() and
 * bench.py's cpu_baseline leg; never by the product.
 *
 * Build: gcc -O2 -fopenmp -shared -fPIC oracle.c -o liboracle.so -lm  (oracle/Makefile)
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* Base table: primality flags for odd q <= BASE_CAP and the odd primes list. */
/* ------------------------------------------------------------------------- */
#define BASE_CAP (1ull << 22) /* covers sqrt(q) for q < 1.76e13 */
#define PRESIEVE_CAP 65536ull /* pre-sieve bound before Miller-Rabin on high windows */

static uint8_t *g_flag;       /* g_flag[q >> 1] = 1 iff odd q is prime, q <= BASE_CAP */
static uint32_t *g_oddp;      /* odd primes <= BASE_CAP, ascending */
static size_t g_noddp;
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void base_init(void) {
    uint64_t n = BASE_CAP;
    uint8_t *mark = (uint8_t *)malloc(n + 1);
    memset(mark, 1, n + 1);
    mark[0] = mark[1] = 0;
    for (uint64_t i = 2; i * i <= n; i++)
        if (mark[i])
            for (uint64_t j = i * i; j <= n; j += i) mark[j] = 0;
    g_flag = (uint8_t *)calloc(n / 2 + 1, 1);
    size_t cnt = 0;
    for (uint64_t q = 3; q <= n; q += 2)
        if (mark[q]) { g_flag[q >> 1] = 1; cnt++; }
    g_oddp = (uint32_t *)malloc(cnt * sizeof(uint32_t));
    g_noddp = 0;
    for (uint64_t q = 3; q <= n; q += 2)
        if (mark[q]) g_oddp[g_noddp++] = (uint32_t)q;
    free(mark);
}

static inline void base_ready(void) { pthread_once(&g_once, base_init); }

/* ------------------------------------------------------------------------- */
uint64_t orc_isqrt(uint64_t n) {
    /* exact floor(sqrt(n)); same contract as the binary search at goldbach.cu:636-646 */
    uint64_t r = (uint64_t)sqrtl((long double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) r--;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) r++;
    return r;
}

/* ------------------------------------------------------------------------- */
/* Deterministic Miller-Rabin, 12 bases 2..37 (goldbach.cu:333-360).          */
/* ------------------------------------------------------------------------- */
static inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

/* n passes the strong test to each of the first k bases 2, 3, 5, ... (odd n > 3) */
int orc_strong_first(uint64_t n, int k) {
    if (n < 2) return 0;
    if (n == 2 || n == 3) return 1;
    if ((n & 1) == 0) return 0;
    uint64_t d = n - 1, r = 0;
    while ((d & 1) == 0) { d >>= 1; r++; }
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (k > 12) k = 12;
    for (int i = 0; i < k; i++) {
        uint64_t a = bases[i];
        if (a >= n) continue;
        uint64_t x = 1, b = a % n, e = d;
        while (e) {
            if (e & 1) x = mulmod(x, b, n);
            b = mulmod(b, b, n);
            e >>= 1;
        }
        if (x == 1 || x == n - 1) continue;
        int ok = 0;
        for (uint64_t j = 0; j + 1 < r; j++) {
            x = mulmod(x, x, n);
            if (x == n - 1) { ok = 1; break; }
        }
        if (!ok) return 0;
    }
    return 1;
}

int orc_is_prime(uint64_t n) { return orc_strong_first(n, 12); }

/* smallest odd composite in [9, limit] passing the first k bases, 0 if none */
uint64_t orc_first_strong_pseudoprime(int k, uint64_t limit) {
    for (uint64_t n = 9; n <= limit; n += 2) {
        if (!orc_strong_first(n, k)) continue;
        for (uint64_t f = 3; f * f <= n; f += 2)
            if (n % f == 0) return n;
    }
    return 0;
}

/* exact primality, table first */
static inline int is_prime_any(uint64_t q) {
    if (q <= BASE_CAP) {
        if (q == 2) return 1;
        if (q < 2 || (q & 1) == 0) return 0;
        return g_flag[q >> 1];
    }
    if ((q & 1) == 0) return 0;
    return orc_is_prime(q);
}

/* ------------------------------------------------------------------------- */
/* simple_sieve / segmented_sieve (segmented_sieve.cpp:13-50).                */
/* ------------------------------------------------------------------------- */
size_t orc_simple_sieve(uint64_t limit, uint64_t *out, size_t cap) {
    if (limit < 2) return 0;
    uint8_t *mark = (uint8_t *)malloc(limit + 1);
    if (!mark) return 0;
    memset(mark, 1, limit + 1);
    mark[0] = mark[1] = 0;
    for (uint64_t i = 2; i * i <= limit; i++)
        if (mark[i])
            for (uint64_t j = i * i; j <= limit; j += i) mark[j] = 0;
    size_t cnt = 0;
    for (uint64_t i = 2; i <= limit; i++)
        if (mark[i]) {
            if (out && cnt < cap) out[cnt] = i;
            cnt++;
        }
    free(mark);
    return cnt;
}

int orc_segmented_sieve(uint64_t low, uint64_t high, uint8_t *out) {
    if (high < low || !out) return -1;
    uint64_t len = high - low + 1;
    memset(out, 1, len);
    if (low == 0) out[0] = 0;
    if (low <= 1 && high >= 1) out[1 - low] = 0;
    uint64_t sq = orc_isqrt(high);
    size_t np = orc_simple_sieve(sq, NULL, 0);
    uint64_t *bp = (uint64_t *)malloc((np ? np : 1) * sizeof(uint64_t));
    if (!bp) return -1;
    orc_simple_sieve(sq, bp, np);
    for (size_t k = 0; k < np; k++) {
        uint64_t p = bp[k];
        /* first multiple of p in [low, high], never p itself (segmented_sieve.cpp:39-43) */
        u128 start = ((u128)low + p - 1) / p * p;
        if (start < 2 * (u128)p) start = 2 * (u128)p;
        for (u128 j = start; j <= high; j += p) out[(uint64_t)j - low] = 0;
    }
    free(bp);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* PrimeBitset (prime_bitset.hpp:29-101, prime_bitset.cpp:21-97).            */
/* ------------------------------------------------------------------------- */
static inline uint64_t num_bits(uint64_t limit) { return limit < 3 ? 0 : (limit - 3) / 2 + 1; }

uint64_t orc_bitset_word_count(uint64_t limit) { return (num_bits(limit) + 63) / 64; }

void orc_prime_bitset(uint64_t limit, uint64_t *words) {
    uint64_t nw = orc_bitset_word_count(limit);
    for (uint64_t w = 0; w < nw; w++) words[w] = ~0ull; /* ctor: all ones (hpp:38) */
    if (limit < 3) return;
    base_ready();
    uint64_t sq = orc_isqrt(limit);
    uint64_t nb = num_bits(limit);
    /* race-free: each thread owns whole 64-bit words */
#pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t w0 = 0; w0 < nw; w0 += 4096) {
        uint64_t w1 = w0 + 4096 < nw ? w0 + 4096 : nw;
        uint64_t lo = 2 * (w0 * 64) + 3;                  /* first odd in slice */
        uint64_t bhi = w1 * 64 < nb ? w1 * 64 : nb;       /* bit end (exclusive) */
        uint64_t hi = 2 * (bhi - 1) + 3;                  /* last odd in slice */
        for (uint64_t p = 3; p <= sq; p += 2) {
            if (!is_prime_any(p)) continue;
            u128 first = (u128)p * p;
            if (first < lo) {
                u128 m = ((u128)lo + p - 1) / p * p;
                if ((m & 1) == 0) m += p;
                first = m;
            }
            for (u128 j = first; j <= hi; j += 2 * (u128)p) {
                uint64_t bit = ((uint64_t)j - 3) / 2;
                words[bit >> 6] &= ~(1ull << (bit & 63));
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Exact odd-only byte window: win[i] = isprime(qlo + 2i), i < nodd.         */
/* ------------------------------------------------------------------------- */
static int build_window(uint64_t qlo, uint64_t qhi, uint8_t *win) {
    uint64_t nodd = (qhi - qlo) / 2 + 1;
    memset(win, 1, nodd);
    uint64_t sq = orc_isqrt(qhi);
    int full = sq <= BASE_CAP;
    uint64_t pmax = full ? sq : PRESIEVE_CAP;
    /* number of table primes used */
    size_t np = 0;
    while (np < g_noddp && g_oddp[np] <= pmax) np++;
    uint64_t *next = (uint64_t *)malloc((np ? np : 1) * sizeof(uint64_t));
    if (!next) return -1;
    for (size_t k = 0; k < np; k++) {
        uint64_t p = g_oddp[k];
        u128 first = (u128)p * p; /* never clear p itself */
        if (first < qlo) {
            u128 m = ((u128)qlo + p - 1) / p * p;
            if ((m & 1) == 0) m += p;
            first = m;
        }
        next[k] = first > qhi ? nodd : (uint64_t)((first - qlo) / 2);
    }
    const uint64_t BLK = 1ull << 18;
    for (uint64_t b0 = 0; b0 < nodd; b0 += BLK) {
        uint64_t b1 = b0 + BLK < nodd ? b0 + BLK : nodd;
        for (size_t k = 0; k < np; k++) {
            uint64_t p = g_oddp[k];
            uint64_t i = next[k];
            for (; i < b1; i += p) win[i] = 0;
            next[k] = i;
        }
    }
    free(next);
    if (!full) {
        /* survivors of the pre-sieve are decided by Miller-Rabin */
        for (uint64_t i = 0; i < nodd; i++)
            if (win[i]) {
                uint64_t q = qlo + 2 * i;
                if (q > PRESIEVE_CAP) win[i] = (uint8_t)orc_is_prime(q);
            }
    }
    return 0;
}

uint64_t orc_window_word_count(uint64_t q_low, uint64_t q_high) {
    if (q_high < q_low) return 0;
    return ((q_high - q_low) / 2 + 1 + 63) / 64;
}

int orc_window_bits(uint64_t q_low, uint64_t q_high, uint64_t *words) {
    /* tiled_sieve_segment_kernel contract (goldbach.cu:198-259): q_low, q_high odd, q_low >= 3 */
    if (q_low < 3 || !(q_low & 1) || !(q_high & 1) || q_high < q_low || !words) return -1;
    base_ready();
    uint64_t nodd = (q_high - q_low) / 2 + 1;
    uint8_t *win = (uint8_t *)malloc(nodd);
    if (!win) return -1;
    if (build_window(q_low, q_high, win)) { free(win); return -1; }
    uint64_t nw = (nodd + 63) / 64;
    for (uint64_t w = 0; w < nw; w++) {
        uint64_t word = ~0ull; /* tail bits stay 1 (tile init at goldbach.cu:219-221) */
        for (uint64_t b = 0; b < 64; b++) {
            uint64_t i = w * 64 + b;
            if (i < nodd && !win[i]) word &= ~(1ull << b);
        }
        words[w] = word;
    }
    free(win);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Minimal partition scan (cpu_goldbach.cpp:19-22 order; single_check's job) */
/* ------------------------------------------------------------------------- */
uint64_t orc_min_partition(uint64_t n, uint64_t bound) {
    if (n < 4 || (n & 1)) return 0;
    base_ready();
    if (n == 4) return bound >= 2 ? 2 : 0;
    uint64_t half = n / 2;
    for (size_t k = 0; k < g_noddp; k++) {
        uint64_t p = g_oddp[k];
        if (p > half || p > bound) return 0;
        if (is_prime_any(n - p)) return p;
    }
    for (uint64_t p = BASE_CAP + 1; p <= half && p <= bound; p += 2)
        if (orc_is_prime(p) && is_prime_any(n - p)) return p;
    return 0;
}

void orc_min_partition_batch(const uint64_t *n, size_t count, uint64_t bound, uint64_t *p_out, int threads) {
    base_ready();
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
    for (size_t i = 0; i < count; i++) p_out[i] = orc_min_partition(n[i], bound);
}

/* ------------------------------------------------------------------------- */
/* One segment: Phase 1 (p <= p_small) + Phase 2 (p <= phase2_limit).        */
/* ------------------------------------------------------------------------- */
#define SEG_HALO 65536ull

static void stats_zero(orc_stats *s) { memset(s, 0, sizeof(*s)); }

static void stats_merge(orc_stats *a, const orc_stats *b) {
    a->evens += b->evens;
    a->phase2 += b->phase2;
    if (b->failures) {
        if (!a->failures || b->first_fail_n < a->first_fail_n) a->first_fail_n = b->first_fail_n;
        a->failures += b->failures;
    }
    a->sum_pmin += b->sum_pmin;
    if (b->max_pmin > a->max_pmin || (b->max_pmin == a->max_pmin && b->max_pmin && b->argmax_n < a->argmax_n)) {
        a->max_pmin = b->max_pmin;
        a->argmax_n = b->argmax_n;
    }
}

int orc_verify_segment(uint64_t A, uint64_t count, uint64_t p_small, uint64_t phase2_limit,
                       uint64_t *verified_bits, uint32_t *pmin_out, orc_stats *st) {
    if (A < 4 || (A & 1) || count == 0) return -1;
    if (count - 1 > (UINT64_MAX - 1 - A) / 2) return -1; /* last n must be <= 2^64-2 */
    base_ready();
    uint64_t B = A + 2 * (count - 1);
    uint64_t bound = p_small > phase2_limit ? p_small : phase2_limit;
    uint64_t qlo = A > SEG_HALO + 3 ? ((A - SEG_HALO) | 1) : 3;
    uint64_t qhi = B + 1;
    uint64_t nodd = (qhi - qlo) / 2 + 1;
    uint8_t *win = (uint8_t *)malloc(nodd);
    if (!win) return -1;
    if (build_window(qlo, qhi, win)) { free(win); return -1; }
    if (verified_bits) memset(verified_bits, 0, ((count + 63) / 64) * sizeof(uint64_t));

    orc_stats s;
    stats_zero(&s);
    s.evens = count;
    for (uint64_t j = 0; j < count; j++) {
        uint64_t n = A + 2 * j;
        uint64_t pm = 0;
        if (n == 4) {
            pm = bound >= 2 ? 2 : 0; /* 2 + 2: the only partition using p = 2 */
        } else {
            uint64_t half = n / 2;
            size_t k = 0;
            for (; k < g_noddp; k++) {
                uint64_t p = g_oddp[k];
                if (p > half || p > bound) break;
                uint64_t q = n - p;
                int pr = q >= qlo ? win[(q - qlo) / 2] : is_prime_any(q);
                if (pr) { pm = p; break; }
            }
            if (!pm && k == g_noddp) {
                for (uint64_t p = BASE_CAP + 1; p <= half && p <= bound; p += 2)
                    if (orc_is_prime(p) && is_prime_any(n - p)) { pm = p; break; }
            }
        }
        int verified = pm && pm <= p_small;
        if (verified_bits && verified) verified_bits[j >> 6] |= 1ull << (j & 63);
        if (pmin_out) pmin_out[j] = (uint32_t)pm;
        if (!verified) s.phase2++;
        if (!pm) {
            if (!s.failures) s.first_fail_n = n;
            s.failures++;
        } else {
            s.sum_pmin += pm;
            if (pm > s.max_pmin) { s.max_pmin = pm; s.argmax_n = n; }
        }
    }
    free(win);
    if (st) *st = s;
    return 0;
}

int orc_verify_range(uint64_t start, uint64_t limit, uint64_t seg_evens, uint64_t p_small,
                     uint64_t phase2_limit, int threads, orc_stats *st) {
    if (limit & 1) limit--;                   /* goldbach.cu:594 */
    if (start < 4) start = 4;                 /* goldbach.cu:604-605 */
    if (start & 1) start++;
    if (limit < 4 || start > limit || seg_evens == 0) return -1;
    uint64_t nseg = ((limit - start) / 2) / seg_evens + 1;
    orc_stats tot;
    stats_zero(&tot);
    int err = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        orc_stats mine;
        stats_zero(&mine);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (uint64_t s = 0; s < nseg; s++) {
            uint64_t a = start + 2 * seg_evens * s;  /* goldbach.cu:422 */
            uint64_t b = limit - a < 2 * seg_evens - 2 ? limit : a + 2 * seg_evens - 2; /* :425 */
            orc_stats one;
            if (orc_verify_segment(a, (b - a) / 2 + 1, p_small, phase2_limit, NULL, NULL, &one)) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
                err = 1;
                continue;
            }
            stats_merge(&mine, &one);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        stats_merge(&tot, &mine);
    }
    if (st) *st = tot;
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* cpu_goldbach (cpu_goldbach.cpp:45-130): byte sieve to LIMIT, then scan.   */
/* ------------------------------------------------------------------------- */
int orc_cpu_goldbach(uint64_t limit, uint64_t *checked, uint64_t *failures, uint64_t *first_fail) {
    if (limit < 4) return -1;
    if (limit & 1) limit--;
    uint8_t *isp = (uint8_t *)malloc(limit + 1);
    if (!isp) return -1;
    orc_segmented_sieve(0, limit, isp);
    uint64_t c = 0, f = 0, ff = 0;
    for (uint64_t n = 4; n <= limit; n += 2) {
        c++;
        int ok = 0;
        for (uint64_t p = 2; p <= n / 2; p++)
            if (isp[p] && isp[n - p]) { ok = 1; break; }
        if (!ok) { f++; ff = n; break; }
    }
    free(isp);
    if (checked) *checked = c;
    if (failures) *failures = f;
    if (first_fail) *first_fail = ff;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Partition counts, cpu_goldbach.cpp:29-36 restated over one byte table.      */
/* ------------------------------------------------------------------------- */
int orc_count_partitions(uint64_t start, uint64_t count, uint64_t *out, int threads) {
    if (start < 4 || (start & 1) || count == 0) return -1;
    const uint64_t n_max = start + 2 * (count - 1);
    uint8_t *isp = (uint8_t *)malloc(n_max + 1);
    if (!isp) return -1;
    if (orc_segmented_sieve(0, n_max, isp) != 0) { free(isp); return -1; }
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        const uint64_t n = start + 2 * (uint64_t)i;
        uint64_t c = 0;
        for (uint64_t p = 2; p <= n / 2; ++p)
            if (isp[p] && isp[n - p]) ++c;
        out[i] = c;
    }
    free(isp);
    return 0;
}
