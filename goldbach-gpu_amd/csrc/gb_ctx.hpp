// gb_ctx.hpp -- internal state of a gb_ctx and the helpers the C-ABI
// translation units share (range verification in gb_api.cpp, the serving
// engine in gb_serve.cpp).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "gb_kernels.hpp"
#include "goldbach_gpu.h"

namespace gbh {

constexpr uint32_t PLAIN_TILE_BITS = 1u << 18;   // plain sieve tile (bits per workgroup)
constexpr uint32_t LANE_ITEM_DIV = 32;            // default: lane items for p * 32 < tile bits (odd-only tiles)
constexpr uint32_t WHEEL_LANE_ITEM_DIV = 128;     // ... and p * 128 < stream bits in the wheel
constexpr uint32_t WAVE_ITEM_DIV = 1024;          // default: wave items for p * 1024 < stream bits
constexpr uint32_t MR_PROOF_STRIDE = 5;           // default: odd-only MR tiles prove one window word in 5

struct ItemTable {
    uint32_t tb = 0;        // tile bits the schedule was cut for
    uint32_t wheel = 0;     // 1: items per (prime, stream) of the mod-6 wheel
    uint32_t n_med = 0;
    uint32_t k_wave = 0;    // wheel: primes below index k_wave are cut into wave items
    uint32_t med_div = 0;   // lane items for p * med_div < tb
    uint32_t wave_div = 0;  // wheel: wave items for p * wave_div < tb
    uint32_t* d_items = nullptr;
    uint32_t* d_items_p = nullptr;
    uint32_t* d_start = nullptr;
};

}  // namespace gbh

struct gb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    gb_config cfg{};
    int cus = 0;
    uint64_t total_mem = 0;
    // device memory the context holds, against its budget (every hipMalloc goes through gbh::dalloc)
    uint64_t budget = 0;
    uint64_t mem_used = 0;
    bool budget_hit = false;
    std::unordered_map<void*, uint64_t> held;
    // prime tables
    uint32_t* d_sp = nullptr;
    double* d_inv = nullptr;
    uint32_t* d_c32 = nullptr;
    uint32_t* d_d5 = nullptr;          // (-4 / 6) mod p per prime (wheel form)
    uint32_t nsp = 0;
    std::vector<uint32_t> h_sp;
    uint64_t* d_p2bits = nullptr;
    uint64_t p2_limit = 0;   // odd
    std::vector<gbh::ItemTable> items;
    // plain segment buffers
    uint64_t* d_window = nullptr;
    uint64_t window_cap = 0;   // words
    uint64_t* d_verified = nullptr;
    uint64_t verified_cap = 0; // words
    uint32_t* d_pmin = nullptr;
    uint64_t pmin_cap = 0;
    gbk::StatRec* d_rec = nullptr;
    uint64_t* d_tinfo = nullptr;   // wheel launches: per tile, window index and sieving-prime count (rec_cap entries)
    uint64_t rec_cap = 0;
    gbk::StatRec* d_out = nullptr;
    uint32_t out_cap = 0;
    // slow list
    uint64_t* d_slow_n = nullptr;
    uint32_t* d_slow_win = nullptr;
    uint64_t* d_slow_ctl = nullptr;   // [0]: count (u32) | overflow (u32) as two words
    uint64_t* d_slow_res = nullptr;
    // windows + reduce ranges
    gbk::WinDesc* d_wins = nullptr;
    uint64_t* d_rec_lo = nullptr;
    uint64_t* d_rec_hi = nullptr;
    uint32_t wins_cap = 0;
    uint64_t* d_counter = nullptr;
    gbk::PrimeRes* d_r0 = nullptr;     // per-window sieving-prime entries
    uint64_t r0_cap = 0;
    uint64_t* d_wcopy = nullptr;       // odd-only launches with sparse proofs: a saved window per workgroup
    uint64_t wcopy_cap = 0;            // words
    // serving buffers
    uint64_t* d_qn = nullptr;
    uint64_t* d_qp = nullptr;
    uint32_t q_cap = 0;
    // host staging (pinned)
    gbk::StatRec* h_out = nullptr;
    gbk::WinDesc* h_wins = nullptr;    // a launch's pieces, copied to d_wins before it runs
    uint64_t* h_lohi = nullptr;        // [rec_lo | rec_hi] of those pieces
    uint64_t* h_ctl = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // plain segments in flight, two deep: each slot's pinned descriptor, result record and
    // completion (the reduction writes both; the host waits for clk->seq == seq, with no event
    // between segments: an event record costs the stream about 5 us)
    struct SegSlot {
        gbk::WinDesc* h_win = nullptr;
        gbk::StatRec* h_out = nullptr;
        gbk::SegClock* clk = nullptr;
        uint64_t seq = 0;                                                      // of the queued segment
        hipEvent_t ev_sieve = nullptr, ev_mark = nullptr, ev_rest = nullptr;   // GB_CFG_KERNEL_TIMES
        uint64_t bits = 0, primes = 0, evens = 0;                              // of the queued segment
    };
    SegSlot seg_slot[2];
    uint64_t seg_seq = 0;        // plain segments queued on this context
    // A second set of plain segment buffers on a stream of its own, taken by a plain range of more
    // than one segment: segment k + 1's sieve then runs beside segment k's marking.  Allocated on
    // first use within the budget; without it (or with GB_CFG_KERNEL_TIMES) segments share one stream.
    struct PlainSet {
        hipStream_t stream = nullptr;
        uint64_t* window = nullptr;
        uint64_t* verified = nullptr;
        gbk::StatRec* rec = nullptr;
        gbk::PrimeRes* r0 = nullptr;
        gbk::WinDesc* wins = nullptr;
        uint64_t* lohi = nullptr;       // [rec_lo | rec_hi]
        gbk::StatRec* out = nullptr;
        uint64_t* slow_n = nullptr;
        uint32_t* slow_win = nullptr;
        uint64_t* slow_res = nullptr;
        uint64_t* slow_ctl = nullptr;
        uint64_t window_cap = 0, verified_cap = 0, rec_cap = 0;   // words, words, records
    } plain2;
    double wall_khz = 0;         // the device wall clock's rate (wall_clock64 ticks per ms)
    void* h_slot_block = nullptr;
    gb_kernel_times ktimes{};
    // batched launches captured as graphs; a graph's kernels read their pieces from d_wins, so
    // one graph serves every chunk of the same shape
    struct Graph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    // autotuner results: key (form, start, limit, p_small, flags) -> tuning
    std::vector<std::pair<std::vector<uint64_t>, gb_tuning>> tune_cache;
    // serving engines running on this context's tables
    std::atomic<int> servers{0};
    // device address of this device's copy of the allocation-free entry point's scratch array
    unsigned char* scratch = nullptr;
};

namespace gbh {

// Records msg as the context's last error and returns code.
int fail(gb_ctx* c, int code, const std::string& msg);

// The reference's range adjustments (goldbach.cu:594,604-609): limit down to even, start up to
// even >= 4; GB_EINVAL when the range is empty.
int normalise_range(gb_ctx* c, uint64_t* start, uint64_t* limit);

// HIP status -> gb_status with a message (device out-of-memory and the context's own budget
// both give GB_ENOMEM).
int hip_fail(gb_ctx* c, hipError_t e, const char* file, int line);

#define HCK(ctx, call)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, __FILE__, __LINE__);                    \
    } while (0)

// Device allocation counted against the context's memory budget (gb_config.mem_budget);
// hipErrorOutOfMemory, without allocating, when it would pass the budget.
template <typename T>
hipError_t dalloc(gb_ctx* c, T** p, uint64_t count) {
    *p = nullptr;
    if (!count) count = 1;
    const uint64_t bytes = count * sizeof(T);
    if (bytes / sizeof(T) != count || c->mem_used + bytes > c->budget) {
        c->budget_hit = true;
        return hipErrorOutOfMemory;
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
    if (e == hipSuccess) {
        c->mem_used += bytes;
        c->held[*p] = bytes;
    }
    return e;
}
void dfree(gb_ctx* c, void* p);

// Medium-prime schedule of the sieve for tiles of at most `tb` bits (cached on c).
// med_div / wave_div: the work-item bounds (0 = LANE_ITEM_DIV / WAVE_ITEM_DIV).
int build_items(gb_ctx* c, uint32_t tb, ItemTable* out, bool wheel = false, uint32_t med_div = 0,
                uint32_t wave_div = 0);
// Device table view of c with the schedule `it`.
gbk::Tables make_tables(const gb_ctx* c, const ItemTable& it);

// p_out[i] = smallest prime p <= bound with n[i] - p prime (0 if none): gb_min_partition_batch
// with a bound other than the context's phase2_limit.
int min_partition_bounded(gb_ctx* c, const uint64_t* n, uint32_t count, uint64_t bound, uint64_t* p_out);

// Destroys the plain form's second stream if the context holds one (its buffers stay): a stream
// kept idle holds one of the process's hardware queues (GPU_MAX_HW_QUEUES, 4 by default), so the
// four lane streams of a server opened next would share queues and queue behind each other.
// The next plain range that takes two segments in flight creates the stream again.
void release_plain2_stream(gb_ctx* c);

}  // namespace gbh
