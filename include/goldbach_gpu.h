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
 *     and q beyond the sieved range is decided by deterministic Miller-Rabin
 *     with the 12 bases 2..37, valid for all 64-bit q).
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
    GB_EAGAIN = 6    /* serving: no result ready / no free slot yet */
} gb_status;

typedef struct gb_ctx gb_ctx;

/* Sizes fixed when the context opens.  Zero selects the default. */
typedef struct gb_config {
    uint64_t max_n;          /* largest even n any call will touch (default 2^64-2) */
    uint64_t max_seg_evens;  /* largest plain segment, in evens (default 2^27) */
    uint64_t phase2_limit;   /* Phase-2 prime bound (goldbach.cu:316, default 1e8) */
    uint64_t mem_budget;     /* device bytes the context may hold (default: all but 1 GiB) */
    uint32_t serve_slots;    /* serving slots (default 65536) */
    uint32_t reserved;
} gb_config;

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
    double   kernel_ms;     /* device time of the call's kernels (HIP events) */
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

/* ---- whole ranges ------------------------------------------------------ */

#define GB_FORM_PLAIN   0u  /* per-segment launches + one host sync per segment */
#define GB_FORM_BATCHED 1u  /* one persistent fused kernel over a device tile queue, hipGraph */

typedef struct gb_tuning {
    uint32_t form;          /* GB_FORM_* */
    uint32_t tile_log2;     /* batched: evens per tile = 2^tile_log2 (0 = autotuned/default) */
    uint64_t seg_evens;     /* plain: evens per segment (0 = 2^27) */
    uint32_t wg_per_cu;     /* batched: persistent workgroups per CU (0 = default) */
    uint32_t flags;         /* GB_WANT_PMIN */
} gb_tuning;

/* All even n in [start, limit] (start even >= 4 after the reference's
 * adjustments, goldbach.cu:594,604-609).  Replaces main's worker spawn and
 * the segment queue (goldbach.cu:689, 761-773, 378-500).  progress_evens
 * (nullable) is updated as work completes, for a progress printer. */
int gb_verify_range(gb_ctx *ctx, uint64_t start, uint64_t limit, uint64_t p_small,
                    const gb_tuning *tuning, gb_seg_stats *stats,
                    volatile uint64_t *progress_evens);

/* ---- uneven high windows (Miller-Rabin for q) -------------------------- */

typedef struct gb_window {
    uint64_t start;   /* even >= 4 */
    uint64_t evens;   /* >= 1; start + 2(evens-1) <= 2^64-2 */
} gb_window;

/* Every window in one persistent launch; per_window (nullable) gets one
 * gb_seg_stats per window, total (nullable) the sum.  Generalises the
 * reference's range loop to windows anywhere below 2^64, where q is decided
 * by a small-prime pre-sieve plus Miller-Rabin (goldbach.cu:160-179). */
int gb_verify_windows(gb_ctx *ctx, const gb_window *windows, uint32_t count, uint64_t p_small,
                      uint32_t flags, gb_seg_stats *per_window, gb_seg_stats *total);

/* ---- single numbers (single_check's job) ------------------------------- */

/* p_out[i] = smallest prime p with n[i]-p prime (n[i] even >= 4), searched
 * up to phase2_limit; 0 if none.  Replaces check_single + goldbach_single_kernel
 * (single_check.cu:95-282), whose p is the first GPU finder, not the minimum. */
int gb_min_partition_batch(gb_ctx *ctx, const uint64_t *n, uint32_t count, uint64_t *p_out);

#ifdef __cplusplus
}
#endif
#endif /* GOLDBACH_GPU_H */
