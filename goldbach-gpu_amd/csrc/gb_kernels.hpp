// gb_kernels.hpp -- host-side view of the HIP kernels in gb_kernels.hip.
// Launch wrappers only; no device code here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gb_device.hpp"

namespace gbk {

constexpr int SIEVE_BS = 512;    // threads per plain-sieve workgroup
constexpr int MARK_BS = 512;     // threads per plain-marking workgroup
constexpr int FUSED_BS = 1024;   // threads per fused persistent workgroup
constexpr int SERVE_WAVES = 4;   // queries per serving workgroup (one wave each)
constexpr uint32_t SERVE_BASES = 7;      // strong-test bases of a 64-bit q (the waves of a wide step's workgroup)
constexpr uint32_t SERVE_WIDE_MAX = 64;  // steps over at most this many slots run one workgroup per slot
constexpr int NMARK = 1027;      // odd primes in [3, 8193]: the primes the window answers

constexpr uint32_t SIEVE_TILE_MAX = 1u << 18;   // plain sieve: window bits per workgroup, at most
constexpr uint32_t SIEVE_WIN_WORDS = SIEVE_TILE_MAX / 32 + 2;
constexpr uint32_t FUSED_TILE_MAX = 1u << 20;   // fused: evens per tile, at most
constexpr uint32_t FUSED_WIN_WORDS = (((FUSED_TILE_MAX + HALO + WIN_PAD) / 32 + 2) + 3) & ~3u;
constexpr uint32_t SIEVE_MRQ = 4096;   // Miller-Rabin queue entries of a plain sieve workgroup
constexpr uint32_t FUSED_MRQ = 7168;   // ... of a fused workgroup (with the window: 157 KB of LDS)

struct MarkParams {
    uint64_t p_small;
    uint64_t bound;       // max(p_small, phase2_limit)
    uint32_t K;           // odd marking primes sp[0..K), all <= min(bound, HALO_P)
    uint32_t fast_final;  // 1 when sp[0..K) reaches bound: unresolved n are failures
    uint32_t flags;       // bit 0: p_min statistics; MP_SPARSE: see proof_stride
    uint32_t proof_stride;   // with MP_SPARSE (odd-only tiles above 2^44): only the survivors in every
                             // proof_stride-th 64-bit word of a tile's window are proven prime; the n
                             // no proven q resolves go to the slow list, which scans them from p = 3
                             // with exact tests
};
constexpr uint32_t MP_SPARSE = 2u;

struct SlowList {
    uint64_t* n;          // unresolved even n (p_min above the window's reach)
    uint32_t* win;        // window index of each entry
    uint32_t* count;      // entries appended (may exceed cap: then *overflow is set)
    uint32_t* overflow;
    uint32_t cap;
    uint32_t pad;
};

// One work window for the fused kernel: `evens` even n from `start`.
struct WinDesc {
    uint64_t start;
    uint64_t evens;
    uint64_t tile0;       // global index of the window's first tile
    uint64_t word0;       // first word of the window in the optional verified map
    uint64_t rbase;       // odd q every tile's q_lo is measured from (q_lo - rbase < 2^52)
    uint64_t r0off;       // first entry of this window's PrimeRes table
    uint32_t np;          // PrimeRes entries held: sieving primes any tile of the window uses
    uint32_t pad;
    int64_t kbase;        // wheel form: k of the first tile's stream bit 0 (q = 6k + 1 or 6k + 5)
};

// A sieving prime as one window sees it: 16 bytes, one load per prime per tile.
struct PrimeRes {
    uint32_t p;
    uint32_t e;     // first bit b >= 0 of the window's base q (rbase) with p | rbase + 2b
    double inv;     // 1/p
};

// A sieving prime as the wheel form's window sees it, 16 bytes like PrimeRes.  The tiles of a
// window lie a whole number t of stream lengths L (= cls_len) above its base, so a tile's first
// S1 mark is (e - t L) mod p = (e - t lp) mod p with lp = L mod p, and t lp mod p comes from
// 32-bit arithmetic: q = floor(t * ratio) in fp32 is off by at most one for t < 2^17
// (ratio = lp / p rounded to fp32: relative error 2^-24), so t lp - q p lies in (-p, 2p) and
// wraps exactly in 32 bits.
struct WheelRes {
    uint32_t p;
    uint32_t e;      // first bit b >= 0 of S1 (q = 6(kbase + b) + 1) with p | q
    uint32_t lp;     // L mod p
    float ratio;     // lp / p
};
static_assert(sizeof(WheelRes) == sizeof(PrimeRes), "wheel entries share the PrimeRes buffer");

// Sieving primes needed for windows whose largest q is q_top.
__host__ __device__ inline uint64_t sieve_prime_bound(uint64_t q_top) {
    uint64_t r = 0;
    // floor(sqrt(q_top)) by bisection on 32-bit candidates (host and device)
    uint64_t lo = 0, hi = 0xFFFFFFFFull;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (mid <= q_top / mid) lo = mid; else hi = mid - 1;
    }
    r = lo;
    return r;
}

// ---- launchers (stream-ordered, no synchronisation) ----

// Odd-only sieve of window bits [0, nbits): bit i <-> q = qv0 + 2(i - lead), i >= lead.
// pr: the window's sieving-prime entries (launch_residues with rbase = qv0).
hipError_t launch_sieve_window(hipStream_t s, const Tables& T, uint64_t qv0, uint32_t lead, uint64_t nbits,
                               uint32_t tile_bits, uint64_t* out, int tail_ones, const PrimeRes* pr);

// pr[wins[w].r0off + k] = sieving prime k as window w sees it, for k < wins[w].np.
hipError_t launch_residues(hipStream_t s, const Tables& T, const WinDesc* wins, uint32_t nwin, uint32_t max_np,
                           PrimeRes* pr);

// Plain marking of one segment from a window already in device memory.
hipError_t launch_mark_segment(hipStream_t s, const Tables& T, const MarkParams& mp, uint64_t n0, uint64_t cnt,
                               const uint64_t* window, uint64_t window_words, uint32_t words_per_block,
                               uint64_t* verified, uint32_t* pmin_out, SlowList sl, StatRec* rec,
                               uint32_t* nrec_out);

// LDS bytes of one fused workgroup for tiles of tile_evens (n_med: Tables::n_med of its schedule).
size_t fused_lds_bytes(uint32_t tile_evens, uint32_t n_med);
constexpr size_t LDS_PER_CU = 160 * 1024;

// Fused persistent kernel over the tiles of `nwin` windows.  With MP_SPARSE, wcopy holds
// FUSED_COPY_WORDS words per workgroup of the grid: each tile saves its pre-sieved window there
// before the strong tests, for the second pass over the n its proven q leave open.
hipError_t launch_fused(hipStream_t s, const Tables& T, const MarkParams& mp, const WinDesc* wins, uint32_t nwin,
                        uint64_t ntiles, uint32_t tile_evens, uint32_t grid, uint64_t* tile_counter,
                        SlowList sl, StatRec* rec, const PrimeRes* pr, uint64_t* wcopy);
constexpr uint32_t FUSED_COPY_WORDS = FUSED_WIN_WORDS / 2;   // 64-bit words of one saved window

// Wheel (mod 6) form of the fused fast path: tiles of 3 * cls_len even n (cls_len per
// class n mod 6), the window kept as two streams of q = 6k + 1 and q = 6k + 5.
constexpr uint32_t WH_HALO = 1408;                   // k below n0 / 6 (covers p <= 8191: 1366)
// n per class of a wheel tile of size index t: 9 * 2^(t - 4), so a tile spans 27 * 2^(t - 3) n.
// At t = 19 two workgroups' streams (2 x 74,096 B) and their dynamic LDS fill a CU's 160 KB;
// a class length that is a multiple of 2^15 gives each of a 512-thread workgroup's threads the
// same number of words to mark.
__host__ __device__ constexpr uint32_t wheel_class_len(uint32_t t) { return 9u << (t - 4); }
constexpr uint32_t WH_CLASS_MAX = wheel_class_len(20);    // n per class per tile, at most (one workgroup per CU)
constexpr uint32_t WH_CLASS_HALF = wheel_class_len(19);   // ... with two half-size workgroups per CU
// u32 per stream for tiles of at most cmax n per class
__host__ __device__ constexpr uint32_t wh_stream_words(uint32_t cmax) { return (cmax + WH_HALO + 64) / 32 + 2; }
size_t wheel_lds_bytes(uint32_t n_med, uint32_t wg_per_cu);
// pr[wins[w].r0off + k], as WheelRes: prime k as window w's stream S1 sees it (e = first stream
// bit of a multiple of p, counted from k = kbase) for tiles of cls_len bits per stream.
hipError_t launch_residues_wheel(hipStream_t s, const Tables& T, const WinDesc* wins, uint32_t nwin,
                                 uint32_t max_np, uint32_t cls_len, PrimeRes* pr);
// The wheel launch: first each tile's window and sieving-prime count into tinfo[ntiles], then
// the persistent wheel kernel over the tiles.
hipError_t launch_wheel(hipStream_t s, const Tables& T, const MarkParams& mp, const WinDesc* wins, uint32_t nwin,
                        uint64_t ntiles, uint32_t cls_len, uint32_t grid, uint32_t wg_per_cu, uint64_t* tile_counter,
                        SlowList sl, StatRec* rec, const PrimeRes* pr, uint64_t* tinfo);

// Resolve slow-list entries: search primes above sp[K-1] up to bound.
hipError_t launch_slow(hipStream_t s, const Tables& T, const MarkParams& mp, SlowList sl, uint64_t* res_p,
                       const WinDesc* wins, uint64_t* verified, uint32_t* pmin_out, uint64_t pmin_n0);

// Merge per-block records of window w in [rec_lo[w], rec_hi[w]) plus its slow entries into out[w]
// (device or pinned host memory); out[w].pad is nonzero when the slow list overflowed.
// Completion of a plain segment, written to pinned host memory by its reduction after the result
// record: the device wall clock (wall_clock64) at the segment's first kernel and at its end, then
// seq behind a system-scope fence, so a host that reads seq sees the record and both times.
struct SegClock {
    uint64_t t0;
    uint64_t t1;
    uint64_t seq;
    uint64_t pad;
};

// One workgroup per window sums the tile records and the slow list's answers into out[w].  With
// clk (one window), thread 0 then writes clk->t0 = *t0, clk->t1 and clk->seq = seq.
hipError_t launch_reduce(hipStream_t s, const MarkParams& mp, const StatRec* rec, const uint64_t* rec_lo,
                         const uint64_t* rec_hi, uint32_t nwin, SlowList sl, const uint64_t* res_p,
                         StatRec* out, SegClock* clk = nullptr, const uint64_t* t0 = nullptr, uint64_t seq = 0);

// One plain segment's set-up: descriptor wins[0] = wd, record bounds [0, nrec), slow-list
// control cleared, and pr[k] for k < wd.np (the window's sieving-prime entries).
// slow_ctl: [slow count, overflow, wall clock at the segment's start] (3 words).
hipError_t launch_segment_prep(hipStream_t s, const Tables& T, const WinDesc& wd, uint64_t nrec, WinDesc* wins,
                               uint64_t* rec_lo, uint64_t* rec_hi, uint64_t* slow_ctl, PrimeRes* pr);

// Zero a set of 64-bit words (graph-safe replacement for hipMemsetAsync).
hipError_t launch_fill(hipStream_t s, uint64_t* p, uint64_t words, uint64_t value);

// Table builders.
hipError_t launch_seed_primes(hipStream_t s, uint32_t limit, uint64_t* bits_out);
hipError_t launch_compact_primes(hipStream_t s, const uint64_t* bits, uint32_t nwords, uint32_t qlimit,
                                 uint32_t* primes_out, uint32_t cap, uint32_t* count_out);
hipError_t launch_prime_aux(hipStream_t s, const uint32_t* sp, uint32_t n, double* inv, uint32_t* c32, uint32_t* d5);

// Partition counts: counts[(n - start)/2] += #{a <= K/2 : P[a] and P[K - a]}, K = (n - 6)/2, for even
// n in [start, n_end]; P: odd prime bitset over indices [0, M] in pwords words (two zero words past M);
// R: pwords words of scratch, filled with P reversed.  counts must start zeroed; n = 4's 2 + 2 is not added.
hipError_t launch_count_partitions(hipStream_t s, const uint64_t* P, uint64_t* R, uint64_t M, uint64_t pwords,
                                   uint64_t start, uint64_t n_end, uint64_t* counts);

// Serving: a slot's device-resident state (n = 0: free) and its per-step answer in pinned host
// memory (state 0: goes on, 1: final answer p, 2: the slot was free).
struct ServeSlot {
    uint64_t n;
    uint64_t k;       // next candidate index of the scan
};
struct ServeDone {
    uint64_t p;       // minimal p (state 1), 0 when none up to the bound
    uint32_t state;
    uint32_t pad;
};

// One step over slots [0, count): adm[i] != 0 starts query adm[i] in slot i (adm and out are
// pinned host memory, slots device memory).  Up to SERVE_WIDE_MAX slots, each slot gets a
// workgroup whose waves test the candidates to different bases (the latency of one strong test
// per chunk); above, a wave per slot tests base 2 and confirms the smallest survivor.
// How a step tells the host it has finished: the workgroups count themselves in *blocks_done
// (device memory, zero between steps) and the last writes seq to *seq_out (pinned host memory).
struct StepSignal {
    uint32_t* blocks_done;
    uint64_t* seq_out;
    uint64_t seq;
};
hipError_t launch_serve_step(hipStream_t s, const Tables& T, const uint64_t* adm, ServeSlot* slots, uint32_t count,
                             uint64_t bound, uint32_t chunks, ServeDone* out, StepSignal sg);

// Serving: p_out[i] = minimal p for ns[i].
hipError_t launch_min_partition(hipStream_t s, const Tables& T, const uint64_t* ns, uint32_t count, uint64_t bound,
                                uint64_t* p_out);

}  // namespace gbk
