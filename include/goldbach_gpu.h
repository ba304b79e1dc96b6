/*
 * goldbach_gpu.h -- C ABI of the MI355X (gfx950) segmented double-sieve
 * Goldbach verifier.  Plain pointers and sizes only; no exceptions, no HIP
 * or C++ types cross this boundary.
 *
 * The reference (isaac-6/goldbach-gpu) has no FFI: its kernels and host
 * driver share one CUDA translation unit.  Each entry point below names the
 * reference code it replaces (file:line) so a driver written against the
 * reference's loop can be pointed here instead.
 *
 * Conventions
 *   - Every function returns a gb_status (0 = GB_OK).  On failure the
 *     message is available from gb_last_error(ctx) until the next call on
 *     that context (gb_last_error(NULL) holds the last gb_open failure).
 *   - The caller owns every host buffer; the context owns all device memory,
 *     its HIP stream and the resident prime tables.
 *   - One context per host thread.  Calls on one context are serialised by
 *     the caller; contexts on one device may be used from different threads.
 *   - Even n, q and p are uint64; no step is approximated (sieves are exact
 *     and q beyond the sieved range is decided by tests that are exact for
 *     every 64-bit q: Miller-Rabin with the prime bases 2, 3, 5, ... up to
 *     the psi_k bound or with Sinclair's seven bases, or Baillie-PSW; all
 *     agree with the reference's 12 bases 2..37).
 *
 * Bit layouts (identical to the reference)
 *   prime bitset  : bit i of the odd-only set <-> 2i+3, LSB-first in uint64
 *                   words (include/prime_bitset.hpp:7-8 of the reference)
 *   window bitset : bit i <-> q_low + 2i, q_low odd (goldbach.cu:173-175);
 *                   bits past the window's last odd are 1 (goldbach.cu:219-221)
 *   verified map  : bit j of word w <-> n = seg_start + 2(64w + j); the
 *                   reference keeps one uint8 per even (goldbach.cu:413),
 *                   bit j set here iff that byte would be 1.  Bits past the
 *                   segment's last even are 0.
 */
#ifndef GOLDBACH_GPU_H
#define GOLDBACH_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gb_status {
    GB_OK = 0,
    GB_EINVAL = 1,   /* argument outside the contract (goldbach.cu:593-622 checks) */
    GB_ENOMEM = 2,   /* device memory guard refused (goldbach.cu:505-541) */
    GB_EDEVICE = 3,  /* HIP runtime error (the reference throws from CUDA_CHECK, goldbach.cu:105-114) */
    GB_ENODEV = 4,   /* no usable device (goldbach.cu:624-626) */
    GB_ERANGE = 5,   /* request beyond a capacity fixed at gb_open */
    GB_EAGAIN = 6,   /* serving: no result ready / no free slot yet */
    GB_EIO = 7       /* run log: file cannot be read or written */
} gb_status;

typedef struct gb_ctx gb_ctx;

/* Sizes fixed when the context opens.  Zero selects the default. */
typedef struct gb_config {
    uint64_t max_n;          /* largest even n any call will touch (default 2^64-2) */
    uint64_t max_seg_evens;  /* largest plain segment, in evens (default 2^27) */
    uint64_t phase2_limit;   /* Phase-2 prime bound (goldbach.cu:316, default 1e8) */
    uint64_t mem_budget;     /* device bytes the context may hold (default: all but 1 GiB); every
                                device allocation of the context counts against it and a call
                                that would pass it returns GB_ENOMEM */
    uint32_t serve_slots;    /* serving slots (default 65536) */
    uint32_t flags;          /* GB_CFG_* (0: none) */
} gb_config;

#define GB_CFG_KERNEL_TIMES 0x1u   /* time the plain form's sieve and marking launches apart (gb_kernel_times_get) */

/* Per segment / per range / per window results. */
typedef struct gb_seg_stats {
    uint64_t evens;         /* even n examined */
    uint64_t phase2;        /* n not verified by Phase 1 (p_min > p_small): the
                               reference's "Phase 2 fallbacks" (goldbach.cu:473) */
    uint64_t failures;      /* n with no prime p <= max(p_small, phase2_limit)
                               giving n-p prime: "Goldbach FAILED" (goldbach.cu:476-480) */
    uint64_t first_fail_n;  /* smallest failing n, 0 if none */
    uint64_t sum_pmin;      /* GB_WANT_PMIN: sum of p_min(n) over n with a partition */
    uint64_t max_pmin;      /* GB_WANT_PMIN: largest p_min(n) */
    uint64_t argmax_n;      /* GB_WANT_PMIN: smallest n with p_min(n) == max_pmin */
    double   kernel_ms;     /* device time of the call's kernels (HIP events; plain segments: the device
                               wall clock from a segment's first kernel to its last) */
} gb_seg_stats;

/* flags for the verify entry points */
#define GB_WANT_PMIN 0x1u   /* also compute sum/max of the minimal partition p_min(n) */

/* ---- context lifecycle ------------------------------------------------- */

/* Opens `device`, builds the resident tables (sieving primes, the odd
 * bitset to phase2_limit) on the device.  Replaces the per-worker set-up of
 * run_gpu_worker (goldbach.cu:388-417) and main's table builds (:661-676). */
int gb_open(int device, const gb_config *cfg, gb_ctx **out);
int gb_close(gb_ctx *ctx);
const char *gb_last_error(const gb_ctx *ctx);

/* Device facts the CLI prints or checks (goldbach.cu:528-539, :624-634). */
int gb_device_count(int *count);
int gb_device_info(int device, char *name, size_t name_len, uint64_t *total_mem, int *compute_units);

/* ---- plain per-segment form (the drop-in) ------------------------------ */

/* Odd-only prime bitset of q in [q_low, q_high] (q_low, q_high odd, 3 <= q_low
 * <= q_high).  bits_out: host buffer of ceil(((q_high-q_low)/2+1)/64) words.
 * Replaces tiled_sieve_segment_kernel (goldbach.cu:198-259) and, with
 * q_low = 3, build_prime_bitset's bit layout (prime_bitset.cpp:21-97). */
int gb_sieve_window(gb_ctx *ctx, uint64_t q_low, uint64_t q_high, uint64_t *bits_out);

/* One segment: all even n in [seg_start, seg_start + 2(seg_even_count-1)].
 * seg_start even >= 4, 1 <= seg_even_count <= max_seg_evens.
 * verified_bits_out (nullable): ceil(seg_even_count/64) words (layout above).
 * pmin_out (nullable): seg_even_count uint32, p_min(n) or 0 when none was
 *   found below max(p_small, phase2_limit).
 * Replaces one pass of run_gpu_worker's loop: sieve, Phase-1 batches,
 * count_unverified and the CPU Phase 2 (goldbach.cu:425-483). */
int gb_verify_segment(gb_ctx *ctx, uint64_t seg_start, uint64_t seg_even_count, uint64_t p_small,
                      uint32_t flags, uint64_t *verified_bits_out, uint32_t *pmin_out,
                      gb_seg_stats *stats);

/* gb_verify_segment without any allocation: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * outputs and errors as gb_verify_segment, with seg_even_count at most 2^20
 * (GB_ERANGE beyond).  Its device memory is one static array per device
 * (9.9 MB, held from the first use of the library on that device to the end
 * of the process) that every call on the device shares: calls that overlap
 * on one GPU, from any context, race and must be serialised by the caller.
 * The call issues no hipMalloc, hipFree or other allocating routine. */
int gb_verify_segment_noalloc(gb_ctx *ctx, uint64_t seg_start, uint64_t seg_even_count, uint64_t p_small,
                              uint32_t flags, uint64_t *verified_bits_out, uint32_t *pmin_out,
                              gb_seg_stats *stats);

/* Device time of the plain form's launches, per kernel, summed over the
 * segments run since the last reset (contexts opened with
 * GB_CFG_KERNEL_TIMES; zeros otherwise).  No counterpart in the reference,
 * which publishes no per-kernel split (benchmarks/bench_5.sh:30-34 parses
 * lines no binary prints).  reset != 0 zeroes the sums after reading. */
typedef struct gb_kernel_times {
    uint64_t segments;      /* plain segments timed */
    uint64_t window_bits;   /* window bits the sieve launches wrote */
    uint64_t sieve_primes;  /* sieving-prime entries (16 B each) they read */
    uint64_t evens;         /* even n the marking launches covered */
    double   sieve_ms;      /* the window sieve (goldbach.cu:198-259's job) */
    double   mark_ms;       /* Phase-1 marking (goldbach.cu:288-311) */
    double   other_ms;      /* the segment's other launches: prime entries, slow path, reduction */
} gb_kernel_times;

int gb_kernel_times_get(gb_ctx *ctx, gb_kernel_times *out, int reset);

/* ---- whole ranges ------------------------------------------------------ */

#define GB_FORM_PLAIN   0u  /* per-segment launches + one host sync per segment */
#define GB_FORM_BATCHED 1u  /* one persistent fused kernel over a device tile queue, hipGraph */

typedef struct gb_tuning {
    uint32_t form;          /* GB_FORM_* */
    uint32_t tile_log2;     /* batched: tile size index t in 12..20 (0 with wg_per_cu 0: 19 with two
                               workgroups per CU for the part at or below 2^44 - 2^24, where tiles
                               use the mod-6 wheel and hold 27 * 2^(t - 4) evens, and 20 with one
                               above, where tiles hold 2^t evens; a value given applies to both parts) */
    uint64_t seg_evens;     /* plain: evens per segment (0 = 2^27) */
    uint32_t wg_per_cu;     /* batched: persistent workgroups per CU (0 = default; 2 needs tile_log2 <= 19) */
    uint32_t flags;         /* GB_WANT_PMIN | GB_TUNE_AUTO */
    uint32_t lane_item_div; /* batched: primes p with p * lane_item_div below a tile's stream bits are
                               cut into work items of 32 consecutive marks, larger ones get a thread
                               each (0 = 128 in the wheel, 32 in odd-only tiles; 8..256) */
    uint32_t wave_item_div; /* batched wheel: primes p with p * wave_item_div below the stream bits
                               are cut into wavefront items of 512 marks (0 = 1024; 512..65536) */
    uint32_t mr_proof_stride; /* batched, odd-only tiles above 2^44 without GB_WANT_PMIN: the strong
                               tests prove the survivors of one 64-bit window word in every s
                               (0 = 5; 1 = every survivor; at most 64); an n that no proven q
                               resolves tries the other pre-sieve survivors q = n - p (p <= 8191)
                               by Baillie-PSW, and what is still open is decided by the exact
                               slow path, so the counts do not depend on s */
} gb_tuning;

#define GB_TUNE_AUTO 0x2u   /* gb_tuning.flags: take tile / segment size from gb_autotune (cached per context) */

/* Times candidate tunings of `form` on sample windows spread over
 * [start, limit] and writes the fastest to *best (its form and flags as
 * given); *evens_per_s (nullable) gets the sample rate it reached.
 *   batched: tile_log2 in 16..20 x wg_per_cu in {1, 2}, where the LDS fits a CU
 *   plain:   seg_evens in 2^24..2^29 (up to max_seg_evens)
 * Ranges too short to sample return the defaults.  The result is cached on
 * the context per (form, start, limit, p_small, flags).  Generalises the
 * reference's offline grid over SEG_SIZE / P_SMALL / BATCH_SIZE
 * (benchmarks/bench_5.sh:10-13) to an on-device search. */
int gb_autotune(gb_ctx *ctx, uint64_t start, uint64_t limit, uint64_t p_small, uint32_t form, uint32_t flags,
                gb_tuning *best, double *evens_per_s);

/* All even n in [start, limit] (start even >= 4 after the reference's
 * adjustments, goldbach.cu:594,604-609).  Replaces main's worker spawn and
 * the segment queue (goldbach.cu:689, 761-773, 378-500).
 *   plain form:   segments of seg_evens, two in flight (on two streams when
 *                 a second buffer set fits the budget), one host wait each.
 *   batched form: the range is split at 2^44 - 2^24 (below, tiles of the
 *                 mod-6 wheel kernel; above, the odd-only fused kernel with
 *                 exact 64-bit tests) and each part is cut into launches of
 *                 at most 2^17 wheel or 2^15 odd-only tiles, so the device
 *                 memory a call holds does not grow with the range (the
 *                 reference's fixed per-segment buffers, goldbach.cu:378-417).
 * Both forms stop after the first segment or launch holding a failure, as
 * the reference stops its workers (goldbach.cu:420,476-480); stats then
 * cover the evens examined up to there.  progress_evens (nullable) is
 * updated after every segment or launch, for a progress printer
 * (goldbach.cu:702-758). */
int gb_verify_range(gb_ctx *ctx, uint64_t start, uint64_t limit, uint64_t p_small,
                    const gb_tuning *tuning, gb_seg_stats *stats,
                    volatile uint64_t *progress_evens);

/* How gb_verify_range cuts [start, limit] with this tuning, without running
 * it (with GB_TUNE_AUTO the autotuner runs first).  Each gb_chunk is one
 * plain segment or one batched launch; *count gets their number and the
 * first cap are written to out (nullable).  No counterpart in the
 * reference, whose cut is fixed (goldbach.cu:422-431). */
#define GB_KERNEL_PLAIN 0u   /* sieve + marking launches of one plain segment */
#define GB_KERNEL_WHEEL 1u   /* persistent mod-6 wheel kernel (every n <= 2^44 - 2^24) */
#define GB_KERNEL_FUSED 2u   /* persistent odd-only kernel (n above that) */

typedef struct gb_chunk {
    uint64_t first_n, last_n;   /* even n covered */
    uint64_t tiles;             /* batched: tiles of the launch (0 for a plain segment) */
    uint32_t kernel;            /* GB_KERNEL_* */
    uint32_t tile_log2;         /* batched: fused tiles of 2^t evens, wheel tiles of 27 * 2^(t-4) */
    uint32_t wg_per_cu;
    uint32_t reserved;
} gb_chunk;

int gb_range_plan(gb_ctx *ctx, uint64_t start, uint64_t limit, uint64_t p_small, const gb_tuning *tuning,
                  gb_chunk *out, uint64_t cap, uint64_t *count);

/* ---- uneven high windows (q by pre-sieve and Baillie-PSW) -------------- */

typedef struct gb_window {
    uint64_t start;   /* even >= 4 */
    uint64_t evens;   /* >= 1; start + 2(evens-1) <= 2^64-2 */
} gb_window;

/* Every window in one persistent launch; per_window (nullable) gets one
 * gb_seg_stats per window, total (nullable) the sum.  Generalises the
 * reference's range loop to windows anywhere below 2^64, where q is decided
 * by a small-prime pre-sieve plus an exact 64-bit test in place of the
 * reference's Miller-Rabin (goldbach.cu:160-179). */
int gb_verify_windows(gb_ctx *ctx, const gb_window *windows, uint32_t count, uint64_t p_small,
                      uint32_t flags, gb_seg_stats *per_window, gb_seg_stats *total);

/* ---- single numbers (single_check's job) ------------------------------- */

/* p_out[i] = smallest prime p with n[i]-p prime (n[i] even >= 4), searched
 * up to phase2_limit; 0 if none.  Replaces check_single + goldbach_single_kernel
 * (single_check.cu:95-282), whose p is the first GPU finder, not the minimum. */
int gb_min_partition_batch(gb_ctx *ctx, const uint64_t *n, uint32_t count, uint64_t *p_out);

/* ---- resumable runs (generalises --start, goldbach.cu:581,604-609,689) -- */

/* The reference resumes a long run by restarting it with --start=N past
 * what the user believes was verified.  A run log records it instead: the
 * range is cut into chunks of chunk_evens even n (0 = 2^36) and every
 * finished chunk appends one CRC-32-checked text line (fsync'd before the
 * next chunk starts).  Calling again with the same log skips the chunks it
 * holds; a torn or corrupt tail is dropped and recomputed; a log written for
 * another (start, limit, p_small, phase2_limit, chunk_evens, GB_WANT_PMIN)
 * is refused with GB_EINVAL.  A chunk with a failure ends the run, as the
 * reference stops at its first failure (goldbach.cu:476-480).
 * max_chunks > 0 stops after that many new chunks (a time-sliced job);
 * *out covers every chunk in the log, its kernel_ms only this call's.  One
 * writer at a time: a log locked by another call returns GB_EAGAIN. */
typedef struct gb_log_header {
    uint64_t start, limit, p_small, phase2_limit, chunk_evens;
    uint32_t want_pmin, reserved;
} gb_log_header;

typedef struct gb_log_record {
    uint64_t chunk, first_n, last_n, evens, phase2, failures, first_fail_n, sum_pmin, max_pmin, argmax_n;
} gb_log_record;

typedef struct gb_log_info {
    uint64_t resumed_chunks;   /* chunks the log already held */
    uint64_t resumed_evens;
    uint64_t done_chunks;      /* chunks in the log after the call */
    uint64_t total_chunks;
    uint32_t complete;         /* the log covers the range, or a failure ended the run */
    uint32_t truncated;        /* a torn or corrupt tail was dropped */
} gb_log_info;

int gb_verify_range_logged(gb_ctx *ctx, uint64_t start, uint64_t limit, uint64_t p_small, const gb_tuning *tuning,
                           const char *log_path, uint64_t chunk_evens, uint64_t max_chunks, gb_seg_stats *out,
                           gb_log_info *info, volatile uint64_t *progress_evens);

/* Reads a run log without a device: its header and the intact, in-order prefix
 * of records (*count of them; the first cap copied to records, which may be
 * NULL).  *torn (nullable) = 1 when bytes follow that prefix.  GB_EIO when the
 * file is missing or empty, GB_EINVAL when its header is damaged. */
int gb_log_read(const char *path, gb_log_header *header, gb_log_record *records, uint64_t cap, uint64_t *count,
                uint32_t *torn);

/* ---- arbitrary-size single n (big_check.cpp:35-144) -------------------- */

typedef struct gb_big_stats {
    uint64_t primes_scanned;   /* odd primes p up to the answer (all of them to p_max if none) */
    uint64_t candidates;       /* of those, p with n - p free of prime factors below 2^20 */
    uint64_t mr_tasks;         /* strong tests run on the device ((p, base) pairs) */
    uint32_t batches;          /* test launches */
    uint32_t words;            /* 32-bit words of n */
    double   kernel_ms;
} gb_big_stats;

/* *p_out = the smallest prime p <= p_max (0: 10^7, big_check.cpp:58) such
 * that n - p is a strong probable prime to each of the first `rounds` prime
 * bases 2, 3, 5, ..., 97 (0: all 25); 0 when there is none.  n is even,
 * >= 4, given as nwords little-endian 32-bit words (at most 2304: 73,728
 * bits).  The reference asks GMP's mpz_probab_prime_p(q, 25)
 * (big_check.cpp:107) for p in batches over threads and reports the first
 * thread's p; here p is the smallest.  Below 2^64 the bases 2..37 decide
 * primality exactly and the call answers through gb_min_partition_batch. */
int gb_big_min_partition(gb_ctx *ctx, const uint32_t *n_words, uint32_t nwords, uint32_t rounds, uint64_t p_max,
                         uint64_t *p_out, gb_big_stats *stats);

/* Decimal digits <-> little-endian 32-bit words (no device needed).
 * from_decimal: GB_EINVAL on an empty string or a non-digit, GB_ERANGE when
 * more than cap words are needed.  to_decimal: GB_ERANGE when cap bytes
 * cannot hold the digits and the terminating zero. */
int gb_big_from_decimal(const char *digits, uint32_t *words, uint32_t cap, uint32_t *nwords);
int gb_big_to_decimal(const uint32_t *words, uint32_t nwords, char *out, size_t cap);

/* ---- partition counts (cpu_goldbach.cpp:29-36, goldbach_count) -------- */

/* counts_out[i] = c(n) for n = start + 2i, i < count: the number of primes
 * p <= n/2 with n - p prime (unordered partitions; 4 = 2 + 2 counts once).
 * start even >= 4; every n at most 2^36.  The reference's goldbach_count
 * scans p from 2 to n/2 over a byte table for one n; here the odd prime
 * bitset and its reversal are built on the device once and each c(n) is an
 * AND + popcount along an anti-diagonal, 64 pairs per word. */
int gb_count_partitions(gb_ctx *ctx, uint64_t start, uint64_t count, uint64_t *counts_out);

/* ---- serving loop with continuous batching ---------------------------- */

/* A serving engine on one context: a host thread that keeps up to `slots`
 * queries resident on the device (their state in device memory, so the
 * device memory granted to the server bounds admission) and runs one step
 * kernel after another over all of them.  Each step advances every resident query by up to
 * `chunks_per_step` chunks of 64 candidate primes; queries that finish leave
 * their slot at the step boundary and queued queries take the free slots for
 * the next step, so arrivals join work already on the device instead of
 * waiting for a batch to drain.  Every result is the minimal p, the same as
 * gb_min_partition_batch gives for that n alone.
 * Replaces single_check's one-number-per-process flow (single_check.cu:136-319). */
typedef struct gb_server gb_server;

typedef struct gb_serve_config {
    uint32_t slots;            /* resident queries per step (0 = the context's serve_slots) */
    uint32_t chunks_per_step;  /* 64-prime chunks per query per step (0 = 4) */
    uint64_t mem_budget;       /* device bytes of query state the server may hold (16 per slot,
                                  allocated from the context's budget); slots shrink to fit
                                  (0 = no limit beyond the context's budget) */
    uint64_t queue_cap;        /* queries waiting for a slot before submit says GB_EAGAIN (0 = 2^22) */
} gb_serve_config;

typedef struct gb_result {
    uint64_t id;         /* submission number, from 0 in submit order */
    uint64_t n;
    uint64_t p;          /* smallest prime p with n-p prime, 0 if none up to phase2_limit */
    uint64_t submit_ns;  /* CLOCK_MONOTONIC when submitted */
    uint64_t done_ns;    /* CLOCK_MONOTONIC when its step completed */
    uint32_t steps;      /* steps the query was resident for */
    uint32_t reserved;
} gb_result;

typedef struct gb_serve_stats {
    uint64_t submitted, completed;
    uint64_t steps;         /* step kernels launched */
    uint64_t step_queries;  /* sum over steps of resident queries */
    uint64_t max_resident;
    uint64_t queued;        /* waiting for a slot now */
    uint64_t resident;      /* on the device now */
    uint32_t slots;         /* slots after the memory budget */
    uint32_t reserved;
    /* host time of the engine threads, summed over them (ns): */
    uint64_t queue_ns;      /* taking queries from the queue and handing results over, lock waits included */
    uint64_t admit_ns;      /* placing admitted queries in slots */
    uint64_t device_ns;     /* from a step's launch until the host sees it finished */
    uint64_t retire_ns;     /* reading the step's answers and building results */
} gb_serve_stats;

/* Starts the engine.  The context must stay open (gb_close refuses with
 * GB_EAGAIN while any of its servers is open); other calls on the context
 * may run while the server does. */
int gb_serve_open(gb_ctx *ctx, const gb_serve_config *cfg, gb_server **out);
/* Queues count even n >= 4 (all or none: GB_EINVAL for an odd n or n < 4,
 * GB_EAGAIN when the queue has no room for all).  first_id (nullable) gets
 * the id of n[0]; the others follow in order.  Thread-safe. */
int gb_serve_submit(gb_server *srv, const uint64_t *n, uint32_t count, uint64_t *first_id);
/* Moves up to max finished results into out, in completion order, waiting up
 * to wait_ms (0: no wait) for the first.  *got = number moved; GB_EDEVICE if
 * the engine stopped on a device error.  Thread-safe. */
int gb_serve_poll(gb_server *srv, gb_result *out, uint32_t max, uint32_t *got, uint32_t wait_ms);
int gb_serve_stats_get(gb_server *srv, gb_serve_stats *st);
/* Stops the engine (queries not yet finished are dropped) and frees it. */
int gb_serve_close(gb_server *srv);
/* The server's last error message; the pointer stays valid until the calling thread's next call. */
const char *gb_serve_last_error(const gb_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* GOLDBACH_GPU_H */
