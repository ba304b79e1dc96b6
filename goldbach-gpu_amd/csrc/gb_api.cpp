// gb_api.cpp -- the C ABI of include/goldbach_gpu.h on top of gb_kernels.hip.
//
// Host orchestration only: every prime test, sieve, mark and reduction runs
// in the HIP kernels.  The host computes launch geometry, builds the static
// medium-prime work schedule from the device-built prime table, moves the
// caller's buffers and maps HIP errors to gb_status.
//
// Reference host code replaced (src/goldbach.cu): run_gpu_worker :378-500,
// validate_hardware_and_limits :505-541, table builds in main :661-676.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "gb_ctx.hpp"

using namespace gbk;
using namespace gbh;

namespace {

constexpr uint64_t DEFAULT_SEG = 1ull << 27;
constexpr uint64_t DEFAULT_P2 = 100000000ull;           // goldbach.cu:316
constexpr uint64_t VRAM_MARGIN = 50ull << 20;            // goldbach.cu:96
constexpr uint32_t MARK_WPB = 2048;                      // plain marking: verified words per workgroup
constexpr uint32_t DEFAULT_TILE_LOG2 = 20;               // fused tile: 2^20 evens
constexpr uint32_t WHEEL_TILE_LOG2 = 19;                 // wheel default: 3 * 9 * 2^15 evens, two workgroups per CU
constexpr uint32_t SLOW_CAP = 1u << 20;
constexpr uint32_t SEED_LIMIT = 4095;                    // trial-division seed (covers sqrt(2^22))
constexpr uint32_t SP_TABLE_CAP = 300000;                // odd primes below 2^22: 295,946
constexpr uint64_t WHEEL_N_MAX = (1ull << 44) - (1ull << 24);   // wheel windows stay where the table sieves q
constexpr uint64_t COUNT_N_MAX = 1ull << 36;             // partition counts: bitsets of 2^35 bits, 4 GiB each

std::mutex g_open_mu;
std::string g_open_err;

// ---- memory of the allocation-free entry point (gb_verify_segment_noalloc) ----
// One segment of at most NA_SEG_MAX evens: window bitset, verified map and p_min staging,
// marking records, the window's sieving-prime entries (every table prime at most), one
// descriptor, the result record and a slow list of NA_SLOW_CAP entries.
constexpr uint64_t NA_SEG_MAX = 1ull << 20;
constexpr uint32_t NA_SLOW_CAP = 1u << 14;
constexpr uint64_t na_align(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t NA_WIN_WORDS = (NA_SEG_MAX + HALO + WIN_PAD + 63) / 64 + 1;
constexpr uint64_t NA_VER_WORDS = (NA_SEG_MAX + 63) / 64;
constexpr uint64_t NA_RECS = (NA_VER_WORDS + MARK_WPB - 1) / MARK_WPB;
constexpr uint64_t NA_OFF_WIN = 0;
constexpr uint64_t NA_OFF_VER = na_align(NA_OFF_WIN + NA_WIN_WORDS * 8);
constexpr uint64_t NA_OFF_PMIN = na_align(NA_OFF_VER + NA_VER_WORDS * 8);
constexpr uint64_t NA_OFF_REC = na_align(NA_OFF_PMIN + NA_SEG_MAX * 4);
constexpr uint64_t NA_OFF_R0 = na_align(NA_OFF_REC + NA_RECS * sizeof(StatRec));
constexpr uint64_t NA_OFF_WINS = na_align(NA_OFF_R0 + (uint64_t)SP_TABLE_CAP * sizeof(PrimeRes));
constexpr uint64_t NA_OFF_LO = na_align(NA_OFF_WINS + sizeof(WinDesc));
constexpr uint64_t NA_OFF_HI = na_align(NA_OFF_LO + 8);
constexpr uint64_t NA_OFF_OUT = na_align(NA_OFF_HI + 8);
constexpr uint64_t NA_OFF_SLOW_N = na_align(NA_OFF_OUT + sizeof(StatRec));
constexpr uint64_t NA_OFF_SLOW_WIN = na_align(NA_OFF_SLOW_N + NA_SLOW_CAP * 8ull);
constexpr uint64_t NA_OFF_SLOW_RES = na_align(NA_OFF_SLOW_WIN + NA_SLOW_CAP * 4ull);
constexpr uint64_t NA_OFF_SLOW_CTL = na_align(NA_OFF_SLOW_RES + NA_SLOW_CAP * 8ull);
constexpr uint64_t NA_SCRATCH_BYTES = na_align(NA_OFF_SLOW_CTL + 24);   // 9.9 MB
static_assert(NA_SCRATCH_BYTES < (12ull << 20), "scratch stays modest");

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch;   // per device: the address of its copy of scratch

void dhostfree(void* p) {
    if (p) (void)hipHostFree(p);
}

uint32_t count_le(const std::vector<uint32_t>& sp, uint64_t x) {
    return (uint32_t)(std::upper_bound(sp.begin(), sp.end(), (uint32_t)std::min<uint64_t>(x, 0xFFFFFFFFull)) - sp.begin());
}

}  // namespace

// Each device holds its own copy, from the first use of the library on it to the end of the
// process.  In host code `scratch` only names the symbol: its device address comes from
// hipGetSymbolAddress, once per device (g_scratch), and only that address reaches kernels.
static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

namespace gbh {

int fail(gb_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(gb_ctx* c, hipError_t e, const char* file, int line) {
    if (e == hipErrorOutOfMemory) {
        char buf[256];
        if (c && c->budget_hit) {
            c->budget_hit = false;
            std::snprintf(buf, sizeof buf, "device memory budget exceeded (%llu MB held of %llu MB)",
                          (unsigned long long)(c->mem_used >> 20), (unsigned long long)(c->budget >> 20));
        } else {
            std::snprintf(buf, sizeof buf, "device out of memory (%llu MB held by the context)",
                          (unsigned long long)(c ? c->mem_used >> 20 : 0));
        }
        return fail(c, GB_ENOMEM, buf);
    }
    return fail(c, GB_EDEVICE, std::string("HIP error at ") + file + ":" + std::to_string(line) + " -- " +
                                   hipGetErrorString(e));
}

void dfree(gb_ctx* c, void* p) {
    if (!p) return;
    (void)hipFree(p);
    auto it = c->held.find(p);
    if (it != c->held.end()) {
        c->mem_used -= it->second;
        c->held.erase(it);
    }
}

// Medium-prime schedule for tiles of at most `tb` bits: prime k with
// p*med_div < tb (med_div = 32 by default, 128 in the wheel) is cut into ceil(tb/(CHUNK_SPAN*p)) items of CHUNK
// marks (an item skips the multiples of 3, so its CHUNK marks span CHUNK_SPAN*p bits); the
// primes above get a thread each.
//
// Wheel form (wheel = true): tb is the bit count of each stream; items are
// encoded (k << 13) | (stream << 12) | chunk and the streams hold no multiples
// of 3.  Primes with p*wave_div < tb (wave_div = 1024 by default: at least 2 * 64 * WAVE_MARKS
// marks per stream) get wave items: ceil(tb/(64*WAVE_MARKS*p)) per stream, one wavefront each,
// lane l marking l, l + 64, ... of the item's 64*WAVE_MARKS marks; the others
// get ceil(tb/(CHUNK*p)) lane items of CHUNK consecutive marks per stream.
int build_items(gb_ctx* c, uint32_t tb, ItemTable* out, bool wheel, uint32_t med_div, uint32_t wave_div) {
    if (!med_div) med_div = wheel ? WHEEL_LANE_ITEM_DIV : LANE_ITEM_DIV;
    if (!wave_div) wave_div = WAVE_ITEM_DIV;
    for (auto& it : c->items)
        if (it.tb == tb && it.wheel == (wheel ? 1u : 0u) && it.med_div == med_div && it.wave_div == wave_div) {
            *out = it;
            return GB_OK;
        }
    ItemTable t;
    t.tb = tb;
    t.wheel = wheel ? 1u : 0u;
    t.med_div = med_div;
    t.wave_div = wave_div;
    std::vector<uint32_t> items, items_p, start;
    uint32_t k = NPAT;
    start.assign(NPAT + 1, 0);
    t.k_wave = NPAT;
    while (k < c->nsp && (uint64_t)c->h_sp[k] * med_div < tb) {
        const uint32_t p = c->h_sp[k];
        const bool wave_item = wheel && (uint64_t)p * wave_div < tb;
        if (wave_item) t.k_wave = k + 1;
        const uint64_t span = (uint64_t)(wave_item ? 64 * WAVE_MARKS : (wheel ? CHUNK : CHUNK_SPAN)) * p;
        const uint32_t chunks = (uint32_t)((tb + span - 1) / span);
        if (chunks > 4096) return fail(c, GB_EINVAL, "tile too large for the medium-prime schedule");
        for (uint32_t st = 0; st < (wheel ? 2u : 1u); ++st)
            for (uint32_t ch = 0; ch < chunks; ++ch) {
                items.push_back(wheel ? ((k << 13) | (st << 12) | ch) : ((k << 12) | ch));
                items_p.push_back(p);
            }
        ++k;
        start.push_back((uint32_t)items.size());
    }
    t.n_med = k;
    // start[k'] for k' in [0, n_med]: items of primes with index < k'
    start.resize(t.n_med + 1);
    HCK(c, dalloc(c, &t.d_items, items.size()));
    HCK(c, dalloc(c, &t.d_items_p, items_p.size()));
    HCK(c, dalloc(c, &t.d_start, start.size()));
    if (!items.empty()) {
        HCK(c, hipMemcpy(t.d_items, items.data(), items.size() * 4, hipMemcpyHostToDevice));
        HCK(c, hipMemcpy(t.d_items_p, items_p.data(), items_p.size() * 4, hipMemcpyHostToDevice));
    }
    HCK(c, hipMemcpy(t.d_start, start.data(), start.size() * 4, hipMemcpyHostToDevice));
    c->items.push_back(t);
    *out = t;
    return GB_OK;
}

Tables make_tables(const gb_ctx* c, const ItemTable& it) {
    Tables T;
    T.sp = c->d_sp;
    T.inv = c->d_inv;
    T.c32 = c->d_c32;
    T.d5 = c->d_d5;
    T.items = it.d_items;
    T.items_p = it.d_items_p;
    T.item_start = it.d_start;
    T.p2bits = c->d_p2bits;
    T.p2_limit = c->p2_limit;
    T.cover = SP_MAX;
    T.nsp = c->nsp;
    T.n_med = it.n_med;
    T.k_wave = it.k_wave;
    return T;
}

void release_plain2_stream(gb_ctx* c) {
    if (!c->plain2.stream) return;
    (void)hipStreamSynchronize(c->plain2.stream);
    (void)hipStreamDestroy(c->plain2.stream);
    c->plain2.stream = nullptr;
}

int normalise_range(gb_ctx* c, uint64_t* start, uint64_t* limit) {
    if (*limit & 1) --*limit;           // goldbach.cu:594
    if (*start < 4) *start = 4;         // goldbach.cu:604-605
    if (*start & 1) ++*start;
    if (*limit < 4) return fail(c, GB_EINVAL, "LIMIT must be >= 4");
    if (*start > *limit) return fail(c, GB_EINVAL, "START must not exceed LIMIT");
    return GB_OK;
}

}  // namespace gbh

namespace {

SlowList slow_list(const gb_ctx* c) {
    SlowList s;
    s.n = c->d_slow_n;
    s.win = c->d_slow_win;
    s.count = reinterpret_cast<uint32_t*>(c->d_slow_ctl);
    s.overflow = reinterpret_cast<uint32_t*>(c->d_slow_ctl) + 2;
    s.cap = SLOW_CAP;
    s.pad = 0;
    return s;
}

MarkParams mark_params(const gb_ctx* c, uint64_t p_small, uint32_t flags) {
    MarkParams mp;
    mp.p_small = p_small;
    mp.bound = std::max(p_small, c->cfg.phase2_limit);
    const uint64_t reach = std::min<uint64_t>(mp.bound, HALO_P);
    mp.K = count_le(c->h_sp, reach);
    mp.fast_final = mp.bound <= HALO_P ? 1u : 0u;
    mp.flags = flags & GB_WANT_PMIN;
    mp.proof_stride = 1;
    return mp;
}

void to_stats(const StatRec& r, double ms, gb_seg_stats* st) {
    if (!st) return;
    st->evens = r.evens;
    st->phase2 = r.phase2;
    st->failures = r.failures;
    st->first_fail_n = r.failures ? r.first_fail : 0;
    st->sum_pmin = r.sum_pmin;
    st->max_pmin = r.max_pmin;
    st->argmax_n = r.max_pmin ? r.argmax : 0;
    st->kernel_ms = ms;
}

void add_stats(gb_seg_stats* a, const gb_seg_stats& b) {
    a->evens += b.evens;
    a->phase2 += b.phase2;
    if (b.failures && (!a->failures || b.first_fail_n < a->first_fail_n)) a->first_fail_n = b.first_fail_n;
    a->failures += b.failures;
    a->sum_pmin += b.sum_pmin;
    if (b.max_pmin > a->max_pmin || (b.max_pmin == a->max_pmin && b.max_pmin && b.argmax_n < a->argmax_n)) {
        a->max_pmin = b.max_pmin;
        a->argmax_n = b.argmax_n;
    }
    a->kernel_ms += b.kernel_ms;
}

int ensure_wcopy(gb_ctx* c, uint64_t words) {
    if (words <= c->wcopy_cap) return GB_OK;
    dfree(c, c->d_wcopy);
    c->d_wcopy = nullptr;
    c->wcopy_cap = 0;
    HCK(c, dalloc(c, &c->d_wcopy, words));
    c->wcopy_cap = words;
    return GB_OK;
}

int ensure_r0(gb_ctx* c, uint64_t n) {
    if (n <= c->r0_cap) return GB_OK;
    dfree(c, c->d_r0);
    c->d_r0 = nullptr;
    c->r0_cap = 0;
    HCK(c, dalloc(c, &c->d_r0, n));
    c->r0_cap = n;
    return GB_OK;
}

// Sieving-prime entries for a window whose first sieved q is rbase and whose
// largest q is q_top: every prime any of its tiles may use.  A tile whose
// sqrt(q) passes the table (2^22) pre-sieves with the whole table and decides
// the survivors by Miller-Rabin.
void set_residues(const gb_ctx* c, WinDesc* wd, uint64_t rbase, uint64_t q_top, uint64_t r0off) {
    const uint64_t s_top = sieve_prime_bound(q_top);
    const uint64_t s_low = sieve_prime_bound(rbase);
    uint64_t bound = s_top;
    if (s_top > SP_MAX) bound = s_low > SP_MAX ? (uint64_t)MR_PRESIEVE : (uint64_t)SP_MAX;
    wd->rbase = rbase;
    wd->r0off = r0off;
    wd->np = count_le(c->h_sp, bound);
    wd->pad = 0;
    wd->kbase = 0;
}

int ensure_wins(gb_ctx* c, uint32_t n) {
    if (n <= c->wins_cap) return GB_OK;
    dfree(c, c->d_wins);
    dfree(c, c->d_rec_lo);
    dfree(c, c->d_rec_hi);
    dfree(c, c->d_out);
    HCK(c, dalloc(c, &c->d_wins, n));
    HCK(c, dalloc(c, &c->d_rec_lo, n));
    HCK(c, dalloc(c, &c->d_rec_hi, n));
    HCK(c, dalloc(c, &c->d_out, n));
    if (c->h_out) dhostfree(c->h_out);
    if (c->h_wins) dhostfree(c->h_wins);
    if (c->h_lohi) dhostfree(c->h_lohi);
    c->h_out = nullptr;
    c->h_wins = nullptr;
    c->h_lohi = nullptr;
    HCK(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_out), n * sizeof(StatRec), hipHostMallocDefault));
    HCK(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_wins), n * sizeof(WinDesc), hipHostMallocDefault));
    HCK(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_lohi), 2ull * n * sizeof(uint64_t), hipHostMallocDefault));
    c->wins_cap = n;
    return GB_OK;
}

int ensure_rec(gb_ctx* c, uint64_t n) {
    if (n <= c->rec_cap) return GB_OK;
    dfree(c, c->d_rec);
    dfree(c, c->d_tinfo);
    c->rec_cap = 0;
    HCK(c, dalloc(c, &c->d_rec, n));
    HCK(c, dalloc(c, &c->d_tinfo, n));
    c->rec_cap = n;
    return GB_OK;
}

// Build the prime tables on the device: trial-division seed to 4095, sieve of
// [3, 2^22) with the seed, compaction, reciprocal tables, then the Phase-2
// bitset of [3, phase2_limit] with the full table.
int build_tables(gb_ctx* c) {
    hipStream_t s = c->stream;
    uint64_t* d_seedbits = nullptr;
    uint32_t* d_seed = nullptr;
    double* d_seed_inv = nullptr;
    uint32_t* d_seed_c32 = nullptr;
    uint32_t* d_cnt = nullptr;
    const uint32_t seed_words = ((SEED_LIMIT - 3) / 2 + 1 + 63) / 64;
    HCK(c, dalloc(c, &d_seedbits, seed_words));
    HCK(c, dalloc(c, &d_seed, 1024));
    HCK(c, dalloc(c, &d_seed_inv, 1024));
    HCK(c, dalloc(c, &d_seed_c32, 1024));
    HCK(c, dalloc(c, &d_cnt, 2));
    HCK(c, launch_seed_primes(s, SEED_LIMIT, d_seedbits));
    HCK(c, launch_compact_primes(s, d_seedbits, seed_words, SEED_LIMIT, d_seed, 1024, d_cnt));
    uint32_t nseed = 0;
    HCK(c, hipMemcpyAsync(&nseed, d_cnt, 4, hipMemcpyDeviceToHost, s));
    HCK(c, hipStreamSynchronize(s));
    if (nseed == 0 || nseed > 1024) return fail(c, GB_EDEVICE, "seed prime count out of range");
    HCK(c, launch_prime_aux(s, d_seed, nseed, d_seed_inv, d_seed_c32, nullptr));

    Tables seedT{};
    seedT.sp = d_seed;
    seedT.inv = d_seed_inv;
    seedT.c32 = d_seed_c32;
    seedT.d5 = nullptr;
    seedT.items = nullptr;
    seedT.items_p = nullptr;
    seedT.item_start = nullptr;
    seedT.p2bits = nullptr;
    seedT.p2_limit = 0;
    seedT.cover = SEED_LIMIT;
    seedT.nsp = nseed;
    seedT.n_med = NPAT;
    seedT.k_wave = NPAT;

    const uint64_t sp_bits = (SP_MAX - 1 - 3) / 2 + 1;   // odd q in [3, 2^22 - 1]
    const uint64_t sp_words = (sp_bits + 63) / 64;
    uint64_t* d_spbits = nullptr;
    HCK(c, dalloc(c, &d_spbits, sp_words));
    // the seed primes as the window [3, 2^22) sees them
    WinDesc* d_w = nullptr;
    PrimeRes* d_pr = nullptr;
    HCK(c, dalloc(c, &d_w, 1));
    HCK(c, dalloc(c, &d_pr, nseed));
    const WinDesc w_seed{3, 0, 0, 0, 3, 0, nseed, 0, 0};
    HCK(c, hipMemcpyAsync(d_w, &w_seed, sizeof w_seed, hipMemcpyHostToDevice, s));
    HCK(c, launch_residues(s, seedT, d_w, 1, nseed, d_pr));
    HCK(c, launch_sieve_window(s, seedT, 3, 0, sp_bits, PLAIN_TILE_BITS, d_spbits, 0, d_pr));
    HCK(c, dalloc(c, &c->d_sp, SP_TABLE_CAP));
    HCK(c, launch_compact_primes(s, d_spbits, (uint32_t)sp_words, SP_MAX - 1, c->d_sp, SP_TABLE_CAP, d_cnt));
    HCK(c, hipMemcpyAsync(&c->nsp, d_cnt, 4, hipMemcpyDeviceToHost, s));
    HCK(c, hipStreamSynchronize(s));
    if (c->nsp == 0 || c->nsp > SP_TABLE_CAP) return fail(c, GB_EDEVICE, "prime table build produced a bad count");
    HCK(c, dalloc(c, &c->d_inv, c->nsp));
    HCK(c, dalloc(c, &c->d_c32, c->nsp));
    HCK(c, dalloc(c, &c->d_d5, c->nsp));
    HCK(c, launch_prime_aux(s, c->d_sp, c->nsp, c->d_inv, c->d_c32, c->d_d5));
    c->h_sp.resize(c->nsp);
    HCK(c, hipMemcpyAsync(c->h_sp.data(), c->d_sp, c->nsp * 4ull, hipMemcpyDeviceToHost, s));
    HCK(c, hipStreamSynchronize(s));
    dfree(c, d_seedbits);
    dfree(c, d_seed);
    dfree(c, d_seed_inv);
    dfree(c, d_seed_c32);
    dfree(c, d_cnt);
    dfree(c, d_spbits);

    // Phase-2 bitset: odd primes <= phase2_limit
    ItemTable it;
    int rc = build_items(c, PLAIN_TILE_BITS, &it);
    if (rc) return rc;
    uint64_t lim = c->cfg.phase2_limit;
    if (lim < 3) lim = 3;
    if (!(lim & 1)) lim -= 1;
    c->p2_limit = lim;
    const uint64_t p2_bits = (lim - 3) / 2 + 1;
    HCK(c, dalloc(c, &c->d_p2bits, (p2_bits + 63) / 64));
    Tables T = make_tables(c, it);
    T.p2bits = nullptr;
    T.p2_limit = 0;
    WinDesc w_p2{3, 0, 0, 0, 0, 0, 0, 0, 0};
    set_residues(c, &w_p2, 3, lim, 0);
    HCK(c, hipStreamSynchronize(s));
    dfree(c, d_pr);
    d_pr = nullptr;
    HCK(c, dalloc(c, &d_pr, w_p2.np));
    HCK(c, hipMemcpyAsync(d_w, &w_p2, sizeof w_p2, hipMemcpyHostToDevice, s));
    HCK(c, launch_residues(s, T, d_w, 1, w_p2.np, d_pr));
    HCK(c, launch_sieve_window(s, T, 3, 0, p2_bits, PLAIN_TILE_BITS, c->d_p2bits, 0, d_pr));
    HCK(c, hipStreamSynchronize(s));
    dfree(c, d_w);
    dfree(c, d_pr);
    return GB_OK;
}

// Device buffers one plain segment runs on: the context's own (grown on demand) or the
// allocation-free entry point's slices of `scratch`.
struct SegBufs {
    uint64_t* window;
    uint64_t window_cap;   // words
    uint64_t* verified;
    uint64_t verified_cap; // words
    uint32_t* pmin;
    uint64_t pmin_cap;
    StatRec* rec;
    uint64_t rec_cap;
    PrimeRes* r0;
    uint64_t r0_cap;
    WinDesc* wins;         // one descriptor
    uint64_t* rec_lo;
    uint64_t* rec_hi;
    StatRec* out;
    SlowList sl;
    uint64_t* slow_res;
    uint64_t* slow_ctl;
};

SegBufs ctx_bufs(const gb_ctx* c) {
    return SegBufs{c->d_window, c->window_cap, c->d_verified, c->verified_cap, c->d_pmin, c->pmin_cap,
                   c->d_rec, c->rec_cap, c->d_r0, c->r0_cap, c->d_wins, c->d_rec_lo, c->d_rec_hi, c->d_out,
                   slow_list(c), c->d_slow_res, c->d_slow_ctl};
}

// Launch geometry of one plain segment: window words, verified words, marking records and the
// window's sieving-prime entries (the descriptor is left in c->h_wins[0]).
struct SegPlan {
    WinGeom g;
    uint64_t wwords, vwords, nblocks;
    uint32_t np;
};

SegPlan plan_segment(gb_ctx* c, uint64_t A, uint64_t cnt, WinDesc* desc) {
    SegPlan sp;
    sp.g = window_geom(A, cnt + HALO + WIN_PAD);
    sp.wwords = ((uint64_t)sp.g.nbits + 63) / 64;
    sp.vwords = (cnt + 63) / 64;
    sp.nblocks = (sp.vwords + MARK_WPB - 1) / MARK_WPB;
    WinDesc wd{A, cnt, 0, 0, 0, 0, 0, 0, 0};
    set_residues(c, &wd, sp.g.qv0, geom_q_top(sp.g), 0);
    *desc = wd;
    sp.np = wd.np;
    return sp;
}

// Queues one plain segment on buffers that are already large enough, staged through slot
// `slot` (its descriptor already written by plan_segment); allocates nothing and does not wait.
int enqueue_segment(gb_ctx* c, hipStream_t s, const SegBufs& B, const SegPlan& sp, uint32_t slot, uint64_t A,
                    uint64_t cnt, const MarkParams& mp, bool want_verified, bool want_pmin) {
    gb_ctx::SegSlot& S = c->seg_slot[slot];
    const ItemTable* it = nullptr;       // the plain tile schedule, built by gb_open
    for (const ItemTable& t : c->items)
        if (t.tb == PLAIN_TILE_BITS && !t.wheel) it = &t;
    if (!it) return fail(c, GB_EDEVICE, "plain sieve schedule missing");
    const Tables T = make_tables(c, *it);
    if (sp.wwords > B.window_cap || sp.vwords > B.verified_cap || (want_pmin && cnt > B.pmin_cap) ||
        sp.nblocks > B.rec_cap || sp.np > B.r0_cap)
        return fail(c, GB_ERANGE, "segment larger than the buffers it runs on");
    const bool timed = (c->cfg.flags & GB_CFG_KERNEL_TIMES) != 0;
    S.seq = ++c->seg_seq;
    S.bits = sp.g.nbits;
    S.primes = sp.np;
    S.evens = cnt;
    // descriptor, record bounds, slow-list reset and prime entries in one launch (no copies)
    HCK(c, launch_segment_prep(s, T, *S.h_win, sp.nblocks, B.wins, B.rec_lo, B.rec_hi, B.slow_ctl, B.r0));
    if (timed) HCK(c, hipEventRecord(S.ev_sieve, s));
    HCK(c, launch_sieve_window(s, T, sp.g.qv0, sp.g.lead, sp.g.nbits, PLAIN_TILE_BITS, B.window, 0, B.r0));
    if (timed) HCK(c, hipEventRecord(S.ev_mark, s));
    uint32_t nrec = 0;
    HCK(c, launch_mark_segment(s, T, mp, A, cnt, B.window, sp.wwords, MARK_WPB, want_verified ? B.verified : nullptr,
                               want_pmin ? B.pmin : nullptr, B.sl, B.rec, &nrec));
    if (timed) HCK(c, hipEventRecord(S.ev_rest, s));
    HCK(c, launch_slow(s, T, mp, B.sl, B.slow_res, B.wins, want_verified ? B.verified : nullptr,
                       want_pmin ? B.pmin : nullptr, A));
    // the record goes straight to the slot's pinned host memory (no read-back copy), then the
    // segment's device times and its sequence number, which the host waits for
    HCK(c, launch_reduce(s, mp, B.rec, B.rec_lo, B.rec_hi, 1, B.sl, B.slow_res, S.h_out, S.clk, B.slow_ctl + 2, S.seq));
    return GB_OK;
}

// Waits for the segment queued in `slot` on stream s: its record to *out, its kernel time to *ms
// and (clk, nullable) the device clock at its start and end.
int finish_segment(gb_ctx* c, hipStream_t s, uint32_t slot, uint32_t slow_cap, StatRec* out, double* ms,
                   uint64_t* clk = nullptr) {
    gb_ctx::SegSlot& S = c->seg_slot[slot];
    // spin on the pinned sequence word; now and then ask the stream, so a device error ends the
    // wait instead of leaving it spinning
    const volatile uint64_t* seq = &S.clk->seq;
    for (uint32_t spin = 1; *seq != S.seq; ++spin) {
        if (spin % 512) continue;
        const hipError_t e = hipStreamQuery(s);
        if (e == hipErrorNotReady) continue;
        HCK(c, e);
        if (*seq != S.seq) return fail(c, GB_EDEVICE, "plain segment ended without its completion record");
    }
    // the record and clocks were written before the sequence word: no read of them may move above
    // the loop that saw it
    std::atomic_thread_fence(std::memory_order_acquire);
    const double f = (double)(S.clk->t1 - S.clk->t0) / c->wall_khz;
    if (ms) *ms = f;
    if (clk) {
        clk[0] = S.clk->t0;
        clk[1] = S.clk->t1;
    }
    if (c->cfg.flags & GB_CFG_KERNEL_TIMES) {
        HCK(c, hipEventSynchronize(S.ev_rest));
        float fs = 0, fm = 0;
        HCK(c, hipEventElapsedTime(&fs, S.ev_sieve, S.ev_mark));
        HCK(c, hipEventElapsedTime(&fm, S.ev_mark, S.ev_rest));
        gb_kernel_times& K = c->ktimes;
        K.segments += 1;
        K.window_bits += S.bits;
        K.sieve_primes += S.primes;
        K.evens += S.evens;
        K.sieve_ms += fs;
        K.mark_ms += fm;
        K.other_ms += f - fs - fm;
    }
    *out = *S.h_out;
    if (out->pad)
        return fail(c, GB_ERANGE, "slow-path list overflow (more than " + std::to_string(slow_cap) +
                                      " n with p_min beyond the window)");
    return GB_OK;
}

// One plain segment on buffers that are already large enough; allocates nothing.
int run_segment_on(gb_ctx* c, const SegBufs& B, uint64_t A, uint64_t cnt, const MarkParams& mp, bool want_verified,
                   bool want_pmin, double* ms, StatRec* out) {
    const SegPlan sp = plan_segment(c, A, cnt, c->seg_slot[0].h_win);
    int rc = enqueue_segment(c, c->stream, B, sp, 0, A, cnt, mp, want_verified, want_pmin);
    if (rc) return rc;
    return finish_segment(c, c->stream, 0, B.sl.cap, out, ms);
}

void free_plain2(gb_ctx* c) {
    gb_ctx::PlainSet& P = c->plain2;
    void* bufs[] = {P.window, P.verified, P.rec, P.r0, P.wins, P.lohi, P.out, P.slow_n, P.slow_win, P.slow_res,
                    P.slow_ctl};
    for (void* b : bufs)
        if (b) dfree(c, b);
    const hipStream_t s = P.stream;
    P = gb_ctx::PlainSet{};
    P.stream = s;
}

// The second set of plain segment buffers for segments of up to cnt evens, allocated or grown
// within the budget.  False (and no error) when it does not fit: the range then runs on one stream.
bool ensure_plain2(gb_ctx* c, uint64_t cnt) {
    gb_ctx::PlainSet& P = c->plain2;
    const uint64_t wwords = (cnt + HALO + WIN_PAD + 63) / 64 + 1, vwords = (cnt + 63) / 64;
    const uint64_t nrec = (vwords + MARK_WPB - 1) / MARK_WPB;
    if (P.window && wwords <= P.window_cap && vwords <= P.verified_cap && nrec <= P.rec_cap) return true;
    free_plain2(c);
    const std::string err = c->err;
    bool ok = dalloc(c, &P.window, wwords) == hipSuccess && dalloc(c, &P.verified, vwords) == hipSuccess &&
              dalloc(c, &P.rec, nrec) == hipSuccess && dalloc(c, &P.r0, c->nsp) == hipSuccess &&
              dalloc(c, &P.wins, 1) == hipSuccess && dalloc(c, &P.lohi, 2) == hipSuccess &&
              dalloc(c, &P.out, 1) == hipSuccess && dalloc(c, &P.slow_n, SLOW_CAP) == hipSuccess &&
              dalloc(c, &P.slow_win, SLOW_CAP) == hipSuccess && dalloc(c, &P.slow_res, SLOW_CAP) == hipSuccess &&
              dalloc(c, &P.slow_ctl, 3) == hipSuccess;
    if (!ok) {
        free_plain2(c);
        c->budget_hit = false;
        c->err = err;
        return false;
    }
    P.window_cap = wwords;
    P.verified_cap = vwords;
    P.rec_cap = nrec;
    return true;
}

SegBufs plain2_bufs(const gb_ctx* c) {
    const gb_ctx::PlainSet& P = c->plain2;
    SlowList sl;
    sl.n = P.slow_n;
    sl.win = P.slow_win;
    sl.count = reinterpret_cast<uint32_t*>(P.slow_ctl);
    sl.overflow = reinterpret_cast<uint32_t*>(P.slow_ctl) + 2;
    sl.cap = SLOW_CAP;
    sl.pad = 0;
    return SegBufs{P.window, P.window_cap, P.verified, P.verified_cap, nullptr, 0, P.rec, P.rec_cap, P.r0, c->nsp,
                   P.wins, P.lohi, P.lohi + 1, P.out, sl, P.slow_res, P.slow_ctl};
}

// The context's buffers grown for plain segments of up to cnt evens (r0 to the whole table, so a
// range's later segments never reallocate it while earlier ones run).
int ensure_plain(gb_ctx* c, uint64_t cnt, bool want_pmin) {
    const uint64_t vwords = (cnt + 63) / 64;
    if ((cnt + HALO + WIN_PAD + 63) / 64 + 1 > c->window_cap || vwords > c->verified_cap)
        return fail(c, GB_ERANGE, "segment larger than max_seg_evens");
    if (want_pmin && cnt > c->pmin_cap) {
        dfree(c, c->d_pmin);
        c->d_pmin = nullptr;
        c->pmin_cap = 0;
        HCK(c, dalloc(c, &c->d_pmin, cnt));
        c->pmin_cap = cnt;
    }
    int rc = ensure_rec(c, (vwords + MARK_WPB - 1) / MARK_WPB);
    if (rc) return rc;
    return ensure_r0(c, c->nsp);
}

// One plain segment already validated, on the context's buffers (grown as needed).
int run_segment(gb_ctx* c, uint64_t A, uint64_t cnt, const MarkParams& mp, bool want_verified, bool want_pmin,
                double* ms, StatRec* out) {
    int rc = ensure_plain(c, cnt, want_pmin);
    if (rc) return rc;
    return run_segment_on(c, ctx_bufs(c), A, cnt, mp, want_verified, want_pmin, ms, out);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* gb_last_error(const gb_ctx* ctx) {
    if (!ctx) return g_open_err.c_str();
    return ctx->err.c_str();
}

int gb_device_count(int* count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return n > 0 ? GB_OK : GB_ENODEV;
}

int gb_device_info(int device, char* name, size_t name_len, uint64_t* total_mem, int* compute_units) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GB_ENODEV;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return GB_EDEVICE;
    if (name && name_len) {
        const char* nm = p.name[0] ? p.name : p.gcnArchName;
        std::snprintf(name, name_len, "%s", nm);
    }
    if (total_mem) *total_mem = p.totalGlobalMem;
    if (compute_units) *compute_units = p.multiProcessorCount;
    return GB_OK;
}

int gb_kernel_times_get(gb_ctx* c, gb_kernel_times* out, int reset) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!out) return fail(c, GB_EINVAL, "gb_kernel_times_get: out is NULL");
    *out = c->ktimes;
    if (reset) c->ktimes = gb_kernel_times{};
    return GB_OK;
}

int gb_open(int device, const gb_config* cfg_in, gb_ctx** out) {
    std::lock_guard<std::mutex> lk(g_open_mu);
    g_open_err.clear();
    if (!out) { g_open_err = "gb_open: out is NULL"; return GB_EINVAL; }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_open_err = "No HIP devices found."; return GB_ENODEV; }
    if (device < 0 || device >= ndev) { g_open_err = "device index out of range"; return GB_ENODEV; }
    gb_ctx* c = new gb_ctx();
    c->device = device;
    if (cfg_in) c->cfg = *cfg_in;
    if (!c->cfg.max_n) c->cfg.max_n = ~0ull - 1;
    if (!c->cfg.max_seg_evens) c->cfg.max_seg_evens = DEFAULT_SEG;
    if (!c->cfg.phase2_limit) c->cfg.phase2_limit = DEFAULT_P2;
    if (!c->cfg.serve_slots) c->cfg.serve_slots = 65536;
    auto bail = [&](int rc) {
        g_open_err = c->err;
        gb_close(c);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) { c->err = "hipSetDevice failed"; return bail(GB_EDEVICE); }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) { c->err = "hipGetDeviceProperties failed"; return bail(GB_EDEVICE); }
    c->cus = p.multiProcessorCount;
    c->total_mem = p.totalGlobalMem;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || wall_khz <= 0) {
        c->err = "hipDeviceGetAttribute(WallClockRate) failed";
        return bail(GB_EDEVICE);
    }
    c->wall_khz = wall_khz;
    if (c->cfg.phase2_limit > (1ull << 36)) { c->err = "phase2_limit above 2^36"; return bail(GB_EINVAL); }

    // device memory guard (validate_hardware_and_limits, goldbach.cu:505-541)
    const uint64_t seg = c->cfg.max_seg_evens;
    const uint64_t window_words = (seg + HALO + WIN_PAD + 63) / 64 + 1;
    const uint64_t verified_words = (seg + 63) / 64;
    const uint64_t p2_bytes = (c->cfg.phase2_limit / 2 / 64 + 2) * 8;
    const uint64_t fixed = 300000ull * 16 + p2_bytes + SLOW_CAP * 28ull + (64ull << 20);
    const uint64_t required = (window_words + verified_words) * 8 + fixed + VRAM_MARGIN;
    const uint64_t budget = c->cfg.mem_budget ? std::min<uint64_t>(c->cfg.mem_budget, c->total_mem)
                                              : c->total_mem - std::min<uint64_t>(c->total_mem / 2, 1ull << 30);
    c->budget = budget;
    if (required > budget) {
        char buf[256];
        std::snprintf(buf, sizeof buf, "insufficient VRAM: required %llu MB, available %llu MB",
                      (unsigned long long)(required >> 20), (unsigned long long)(budget >> 20));
        c->err = buf;
        return bail(GB_ENOMEM);
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) { c->err = "hipStreamCreate failed"; return bail(GB_EDEVICE); }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        c->err = "hipEventCreate failed";
        return bail(GB_EDEVICE);
    }
    int rc = build_tables(c);
    if (rc) return bail(rc);
    {
        // the scratch array's address on this device (current above), looked up once per device
        std::lock_guard<std::mutex> sk(g_scratch_mu);
        if (g_scratch.size() <= (size_t)device) g_scratch.resize(device + 1, nullptr);
        if (!g_scratch[device]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                c->err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
                return bail(GB_EDEVICE);
            }
            g_scratch[device] = static_cast<unsigned char*>(p);
        }
        c->scratch = g_scratch[device];
    }
    if (dalloc(c, &c->d_window, window_words) != hipSuccess || dalloc(c, &c->d_verified, verified_words) != hipSuccess ||
        dalloc(c, &c->d_slow_n, SLOW_CAP) != hipSuccess || dalloc(c, &c->d_slow_win, SLOW_CAP) != hipSuccess ||
        dalloc(c, &c->d_slow_res, SLOW_CAP) != hipSuccess || dalloc(c, &c->d_slow_ctl, 3) != hipSuccess ||
        dalloc(c, &c->d_counter, 1) != hipSuccess) {
        c->err = "device allocation failed";
        return bail(GB_ENOMEM);
    }
    c->window_cap = window_words;
    c->verified_cap = verified_words;
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_ctl), 16, hipHostMallocDefault) != hipSuccess) {
        c->err = "pinned allocation failed";
        return bail(GB_ENOMEM);
    }
    {
        // per-slot staging of the plain segments in flight: 256 pinned bytes per slot
        constexpr size_t SLOT = 256;
        static_assert(sizeof(WinDesc) + sizeof(StatRec) + sizeof(SegClock) <= SLOT, "slot staging fits");
        if (hipHostMalloc(&c->h_slot_block, 2 * SLOT, hipHostMallocDefault) != hipSuccess) {
            c->err = "pinned allocation failed";
            return bail(GB_ENOMEM);
        }
        for (int k = 0; k < 2; ++k) {
            unsigned char* b = static_cast<unsigned char*>(c->h_slot_block) + k * SLOT;
            gb_ctx::SegSlot& S = c->seg_slot[k];
            S.h_win = reinterpret_cast<WinDesc*>(b);
            S.h_out = reinterpret_cast<StatRec*>(b + sizeof(WinDesc));
            S.clk = reinterpret_cast<SegClock*>(b + sizeof(WinDesc) + sizeof(StatRec));
            S.clk->seq = 0;
            if (hipEventCreate(&S.ev_sieve) != hipSuccess || hipEventCreate(&S.ev_mark) != hipSuccess ||
                hipEventCreate(&S.ev_rest) != hipSuccess) {
                c->err = "hipEventCreate failed";
                return bail(GB_EDEVICE);
            }
        }
    }
    rc = ensure_wins(c, 64);
    if (rc) return bail(rc);
    *out = c;
    return GB_OK;
}

int gb_close(gb_ctx* c) {
    if (!c) return GB_OK;
    if (c->servers.load() > 0) return fail(c, GB_EAGAIN, "gb_close: close the context's servers first");
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->plain2.stream) (void)hipStreamSynchronize(c->plain2.stream);
    for (auto& g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto& it : c->items) {
        dfree(c, it.d_items);
        dfree(c, it.d_items_p);
        dfree(c, it.d_start);
    }
    void* bufs[] = {c->d_sp, c->d_inv, c->d_c32, c->d_d5, c->d_p2bits, c->d_window, c->d_verified, c->d_pmin, c->d_rec,
                    c->d_out, c->d_slow_n, c->d_slow_win, c->d_slow_ctl, c->d_slow_res, c->d_wins, c->d_rec_lo,
                    c->d_rec_hi, c->d_counter, c->d_qn, c->d_qp, c->d_r0, c->d_tinfo, c->d_wcopy};
    for (void* b : bufs)
        if (b) dfree(c, b);
    free_plain2(c);
    if (c->plain2.stream) (void)hipStreamDestroy(c->plain2.stream);
    for (auto& kv : c->held) (void)hipFree(kv.first);   // anything a failed call left behind
    c->held.clear();
    if (c->h_out) dhostfree(c->h_out);
    if (c->h_wins) dhostfree(c->h_wins);
    if (c->h_lohi) dhostfree(c->h_lohi);
    if (c->h_ctl) dhostfree(c->h_ctl);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (auto& S : c->seg_slot)
        for (hipEvent_t e : {S.ev_sieve, S.ev_mark, S.ev_rest})
            if (e) (void)hipEventDestroy(e);
    if (c->h_slot_block) dhostfree(c->h_slot_block);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GB_OK;
}

int gb_sieve_window(gb_ctx* c, uint64_t q_low, uint64_t q_high, uint64_t* bits_out) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!bits_out) return fail(c, GB_EINVAL, "bits_out is NULL");
    if (q_low < 3 || !(q_low & 1) || !(q_high & 1) || q_high < q_low)
        return fail(c, GB_EINVAL, "window must satisfy 3 <= q_low <= q_high, both odd");
    HCK(c, hipSetDevice(c->device));
    ItemTable it;
    int rc = build_items(c, PLAIN_TILE_BITS, &it);
    if (rc) return rc;
    const Tables T = make_tables(c, it);
    const uint64_t nbits = (q_high - q_low) / 2 + 1;
    // stream the window through the segment buffer in word-aligned chunks
    const uint64_t chunk_bits = (c->window_cap - 1) * 64;
    rc = ensure_wins(c, 1);
    if (rc) return rc;
    for (uint64_t b0 = 0; b0 < nbits; b0 += chunk_bits) {
        const uint64_t nb = std::min(chunk_bits, nbits - b0);
        const uint64_t words = (nb + 63) / 64;
        const uint64_t qlo = q_low + 2 * b0;
        WinDesc wd{0, 0, 0, 0, 0, 0, 0, 0, 0};
        set_residues(c, &wd, qlo, qlo + 2 * (nb - 1), 0);
        rc = ensure_r0(c, wd.np);
        if (rc) return rc;
        HCK(c, hipMemcpyAsync(c->d_wins, &wd, sizeof(wd), hipMemcpyHostToDevice, c->stream));
        HCK(c, launch_residues(c->stream, T, c->d_wins, 1, wd.np, c->d_r0));
        HCK(c, launch_sieve_window(c->stream, T, qlo, 0, nb, PLAIN_TILE_BITS, c->d_window,
                                   b0 + nb == nbits ? 1 : 0, c->d_r0));
        HCK(c, hipMemcpyAsync(bits_out + b0 / 64, c->d_window, words * 8, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
    }
    return GB_OK;
}

int gb_verify_segment(gb_ctx* c, uint64_t A, uint64_t cnt, uint64_t p_small, uint32_t flags, uint64_t* verified_out,
                      uint32_t* pmin_out, gb_seg_stats* st) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (A < 4 || (A & 1)) return fail(c, GB_EINVAL, "seg_start must be even and >= 4");
    if (cnt == 0) return fail(c, GB_EINVAL, "seg_even_count must be >= 1");
    if (cnt - 1 > (~0ull - 1 - A) / 2) return fail(c, GB_EINVAL, "segment runs past 2^64-2");
    if (cnt > c->cfg.max_seg_evens) return fail(c, GB_ERANGE, "seg_even_count above max_seg_evens");
    HCK(c, hipSetDevice(c->device));
    const MarkParams mp = mark_params(c, p_small, flags | (pmin_out ? GB_WANT_PMIN : 0));
    double ms = 0;
    StatRec rec;
    int rc = run_segment(c, A, cnt, mp, verified_out != nullptr, pmin_out != nullptr, &ms, &rec);
    if (rc) return rc;
    if (verified_out)
        HCK(c, hipMemcpy(verified_out, c->d_verified, ((cnt + 63) / 64) * 8, hipMemcpyDeviceToHost));
    if (pmin_out) HCK(c, hipMemcpy(pmin_out, c->d_pmin, cnt * 4, hipMemcpyDeviceToHost));
    to_stats(rec, ms, st);
    return GB_OK;
}

int gb_verify_segment_noalloc(gb_ctx* c, uint64_t A, uint64_t cnt, uint64_t p_small, uint32_t flags,
                              uint64_t* verified_out, uint32_t* pmin_out, gb_seg_stats* st) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (A < 4 || (A & 1)) return fail(c, GB_EINVAL, "seg_start must be even and >= 4");
    if (cnt == 0) return fail(c, GB_EINVAL, "seg_even_count must be >= 1");
    if (cnt - 1 > (~0ull - 1 - A) / 2) return fail(c, GB_EINVAL, "segment runs past 2^64-2");
    if (cnt > NA_SEG_MAX)
        return fail(c, GB_ERANGE, "seg_even_count above 2^20, the most gb_verify_segment_noalloc's scratch holds");
    if (!c->scratch) return fail(c, GB_EDEVICE, "scratch address not set up");
    HCK(c, hipSetDevice(c->device));
    unsigned char* base = c->scratch;
    SegBufs B;
    B.window = reinterpret_cast<uint64_t*>(base + NA_OFF_WIN);
    B.window_cap = NA_WIN_WORDS;
    B.verified = reinterpret_cast<uint64_t*>(base + NA_OFF_VER);
    B.verified_cap = NA_VER_WORDS;
    B.pmin = reinterpret_cast<uint32_t*>(base + NA_OFF_PMIN);
    B.pmin_cap = NA_SEG_MAX;
    B.rec = reinterpret_cast<StatRec*>(base + NA_OFF_REC);
    B.rec_cap = NA_RECS;
    B.r0 = reinterpret_cast<PrimeRes*>(base + NA_OFF_R0);
    B.r0_cap = SP_TABLE_CAP;
    B.wins = reinterpret_cast<WinDesc*>(base + NA_OFF_WINS);
    B.rec_lo = reinterpret_cast<uint64_t*>(base + NA_OFF_LO);
    B.rec_hi = reinterpret_cast<uint64_t*>(base + NA_OFF_HI);
    B.out = reinterpret_cast<StatRec*>(base + NA_OFF_OUT);
    B.slow_ctl = reinterpret_cast<uint64_t*>(base + NA_OFF_SLOW_CTL);
    B.slow_res = reinterpret_cast<uint64_t*>(base + NA_OFF_SLOW_RES);
    B.sl.n = reinterpret_cast<uint64_t*>(base + NA_OFF_SLOW_N);
    B.sl.win = reinterpret_cast<uint32_t*>(base + NA_OFF_SLOW_WIN);
    B.sl.count = reinterpret_cast<uint32_t*>(B.slow_ctl);
    B.sl.overflow = reinterpret_cast<uint32_t*>(B.slow_ctl) + 2;
    B.sl.cap = NA_SLOW_CAP;
    B.sl.pad = 0;
    const MarkParams mp = mark_params(c, p_small, flags | (pmin_out ? GB_WANT_PMIN : 0));
    double ms = 0;
    StatRec rec;
    int rc = run_segment_on(c, B, A, cnt, mp, verified_out != nullptr, pmin_out != nullptr, &ms, &rec);
    if (rc) return rc;
    if (verified_out)
        HCK(c, hipMemcpyAsync(verified_out, B.verified, ((cnt + 63) / 64) * 8, hipMemcpyDeviceToHost, c->stream));
    if (pmin_out) HCK(c, hipMemcpyAsync(pmin_out, B.pmin, cnt * 4, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    to_stats(rec, ms, st);
    return GB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched form.  A range (or a list of windows) is cut into pieces on the tile
// grid of its kernel kind: the mod-6 wheel where every n is at most
// WHEEL_N_MAX, the odd-only fused kernel above.  Pieces run in launches of at
// most CHUNK_TILES tiles, so the per-tile records, prime entries and window
// descriptors a launch needs stay bounded whatever the range's length (the
// reference's fixed per-segment buffers, goldbach.cu:378-417).  A launch is a
// graph whose kernels read their pieces from d_wins, so every chunk of the
// same shape replays one graph; the host syncs once per launch, which is where
// progress moves and a failure stops the range (goldbach.cu:420,476-480).
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t CHUNK_TILES = 1ull << 17;   // tiles per launch at most: 8 MiB of records
// Odd-only launches hold at most 2^15 tiles of 2^20 evens: their sparse proofs leave about 1e-5
// of the n open, well within one launch's slow list of 2^20 (an overflowing launch runs again).
constexpr uint64_t FUSED_CHUNK_TILES = 1ull << 15;
constexpr uint32_t LAUNCH_PIECES = 64;          // pieces per launch at most
constexpr size_t GRAPH_CACHE = 8;

struct Shape {
    bool wheel;
    uint32_t tile_log2;   // fused: tiles of 2^tile_log2 evens; wheel: 3 * wheel_class_len(tile_log2) evens
    uint32_t wg;          // persistent workgroups per CU
    uint32_t med_div = 0;    // work-item bounds of the medium-prime schedule (0 = defaults)
    uint32_t wave_div = 0;
    uint32_t proof_stride = 0;   // odd-only MR tiles: survivors proven in one word in this many (0 = 4)
};

// Even n [start, start + 2(evens - 1)] that belong to output `owner`.
struct Piece {
    uint64_t start, evens;
    uint32_t owner;
};

// Tile shape of a kind: the defaults when neither knob is given, else the knobs (one workgroup per
// CU and 2^20-even tiles where only one of them is).
Shape shape_of(bool wheel, uint32_t tile_log2, uint32_t wg, uint32_t med_div = 0, uint32_t wave_div = 0,
               uint32_t proof_stride = 0) {
    if (tile_log2 == 0 && wg == 0)
        return wheel ? Shape{true, WHEEL_TILE_LOG2, 2, med_div, wave_div, proof_stride}
                     : Shape{false, DEFAULT_TILE_LOG2, 1, med_div, wave_div, proof_stride};
    return Shape{wheel, tile_log2 ? tile_log2 : DEFAULT_TILE_LOG2, wg ? wg : 1, med_div, wave_div, proof_stride};
}

int check_shape(gb_ctx* c, const Shape& sh) {
    if (sh.tile_log2 < 12 || sh.tile_log2 > 20) return fail(c, GB_EINVAL, "tile_log2 must be in [12, 20]");
    if (sh.wg > 16) return fail(c, GB_EINVAL, "wg_per_cu must be at most 16");
    if (sh.wheel && sh.wg >= 2 && wheel_class_len(sh.tile_log2) > WH_CLASS_HALF)
        return fail(c, GB_EINVAL, "two workgroups per CU take tiles of at most 2^19 evens");
    if (sh.med_div && (sh.med_div < 8 || sh.med_div > 256)) return fail(c, GB_EINVAL, "lane_item_div must be in [8, 256]");
    if (sh.wave_div && (sh.wave_div < 512 || sh.wave_div > 65536))
        return fail(c, GB_EINVAL, "wave_item_div must be in [512, 65536]");
    if (sh.proof_stride > 64) return fail(c, GB_EINVAL, "mr_proof_stride must be at most 64");
    return GB_OK;
}

uint64_t launch_tiles(const Shape& sh) { return sh.wheel ? CHUNK_TILES : FUSED_CHUNK_TILES; }

uint64_t tile_span(const Shape& sh) {   // n covered by one tile
    return sh.wheel ? 6ull * wheel_class_len(sh.tile_log2) : (2ull << sh.tile_log2);
}

uint64_t piece_tiles(const Shape& sh, uint64_t start, uint64_t evens) {
    const uint64_t n_hi = start + 2 * (evens - 1);
    if (sh.wheel) return (n_hi - (start - start % 6)) / tile_span(sh) + 1;
    const uint64_t te = 1ull << sh.tile_log2;
    return (evens + te - 1) / te;
}

// [a, b] (even, a <= b) cut into pieces of at most cap tiles on the tile grid of sh that starts at
// a (a wheel tile begins at a multiple of 6, so every later piece starts on its own tile).
void cut_pieces(const Shape& sh, uint64_t a, uint64_t b, uint32_t owner, uint64_t cap, std::vector<Piece>& out) {
    const uint64_t nt = piece_tiles(sh, a, (b - a) / 2 + 1);
    const uint64_t base = sh.wheel ? a - a % 6 : a, span = tile_span(sh);
    for (uint64_t t0 = 0; t0 < nt; t0 += cap) {
        const uint64_t t1 = std::min(nt, t0 + cap);
        const uint64_t lo = std::max(a, base + span * t0);
        const uint64_t hi = t1 == nt ? b : base + span * t1 - 2;
        out.push_back(Piece{lo, (hi - lo) / 2 + 1, owner});
    }
}

// [a, b] split where the wheel stops serving (every n of a wheel part is at most WHEEL_N_MAX).
void split_kinds(uint64_t a, uint64_t b, std::vector<std::pair<bool, std::pair<uint64_t, uint64_t>>>& out) {
    if (a <= WHEEL_N_MAX) out.push_back({true, {a, std::min(b, WHEEL_N_MAX)}});
    if (b > WHEEL_N_MAX) out.push_back({false, {std::max(a, WHEEL_N_MAX + 2), b}});
}

void drop_graphs(gb_ctx* c) {
    for (auto& g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

// One launch over pieces pcs[0 .. np) of shape sh; the per-piece records land in c->h_out[0 .. np).
// *overflow: more n than the slow list holds had p_min beyond the window's reach (the launch's
// results are then incomplete and it must run again in smaller pieces).
int run_launch(gb_ctx* c, const Shape& sh, const Piece* pcs, uint32_t np, uint64_t p_small, uint32_t flags,
               double* ms, bool* overflow) {
    hipStream_t s = c->stream;
    MarkParams mp = mark_params(c, p_small, flags);
    // odd-only tiles with the fast marking: their strong tests may prove a subset of the survivors
    // (the slow list then decides the n left open exactly)
    const uint32_t stride = sh.proof_stride ? sh.proof_stride : MR_PROOF_STRIDE;
    if (!sh.wheel && stride > 1 && mp.K == NMARK && mp.p_small >= HALO_P && !mp.fast_final &&
        !(mp.flags & GB_WANT_PMIN)) {
        mp.flags |= MP_SPARSE;
        mp.proof_stride = stride;
    }
    const uint32_t cls_len = wheel_class_len(sh.tile_log2), te = 1u << sh.tile_log2;
    ItemTable it;
    int rc = sh.wheel ? build_items(c, cls_len + WH_HALO + 64, &it, true, sh.med_div, sh.wave_div)
                      : build_items(c, te + HALO + WIN_PAD, &it, false, sh.med_div);
    if (rc) return rc;
    const size_t lds = sh.wheel ? wheel_lds_bytes(it.n_med, sh.wg) : fused_lds_bytes(te, it.n_med);
    // (a schedule whose LDS lets fewer workgroups share a CU than asked runs correctly, with the
    // workgroups that do not fit starting as others finish; gb_autotune skips such shapes)
    if (lds > LDS_PER_CU)
        return fail(c, GB_EINVAL, "the work items of this lane_item_div need more LDS than a CU holds at this tile size");
    const Tables T = make_tables(c, it);
    rc = ensure_wins(c, np);
    if (rc) return rc;
    uint64_t tiles = 0, nres = 0;
    for (uint32_t w = 0; w < np; ++w) {
        WinDesc& wd = c->h_wins[w];
        wd = WinDesc{pcs[w].start, pcs[w].evens, 0, 0, 0, 0, 0, 0, 0};
        const uint64_t nt = piece_tiles(sh, wd.start, wd.evens);
        if (sh.wheel && nt > CHUNK_TILES) return fail(c, GB_EINVAL, "wheel piece longer than a launch's tiles");
        wd.tile0 = tiles;
        c->h_lohi[w] = tiles;
        tiles += nt;
        c->h_lohi[np + w] = tiles;
        if (sh.wheel) {
            const uint64_t nb = wd.start - wd.start % 6;
            wd.kbase = (int64_t)(nb / 6) - (int64_t)WH_HALO;
            const int64_t k_last = wd.kbase + (int64_t)((uint64_t)cls_len * (nt - 1));
            const uint64_t q_top = 6ull * (uint64_t)(k_last + cls_len + WH_HALO + 64 - 1) + 5;
            wd.r0off = nres;
            wd.np = count_le(c->h_sp, sieve_prime_bound(q_top));
        } else {
            // every tile's first sieved q is >= rbase; the last tile reaches q_top
            const uint64_t rbase = wd.start >= 2ull * HALO + 4 ? wd.start - 2ull * HALO - 1 : 3;
            const uint64_t last_n0 = wd.start + 2ull * te * ((wd.evens - 1) / te);
            const uint64_t last_cnt = wd.evens - (uint64_t)te * ((wd.evens - 1) / te);
            const WinGeom lg = window_geom(last_n0, last_cnt + HALO + WIN_PAD);
            set_residues(c, &wd, rbase, geom_q_top(lg), nres);
        }
        nres += wd.np;
    }
    uint64_t grid = (uint64_t)c->cus * sh.wg;
    if (grid > tiles) grid = tiles;
    const void* old_bufs[] = {c->d_rec, c->d_r0, c->d_tinfo, c->d_wcopy};
    rc = ensure_rec(c, tiles);
    if (rc) return rc;
    rc = ensure_r0(c, std::max<uint64_t>(nres, c->nsp));
    if (rc) return rc;
    if (mp.flags & MP_SPARSE) {
        rc = ensure_wcopy(c, grid * FUSED_COPY_WORDS);
        if (rc) return rc;
    }
    if (old_bufs[0] != c->d_rec || old_bufs[1] != c->d_r0 || old_bufs[2] != c->d_tinfo || old_bufs[3] != c->d_wcopy)
        drop_graphs(c);
    HCK(c, hipMemcpyAsync(c->d_wins, c->h_wins, np * sizeof(WinDesc), hipMemcpyHostToDevice, s));
    HCK(c, hipMemcpyAsync(c->d_rec_lo, c->h_lohi, np * 8ull, hipMemcpyHostToDevice, s));
    HCK(c, hipMemcpyAsync(c->d_rec_hi, c->h_lohi + np, np * 8ull, hipMemcpyHostToDevice, s));
    SlowList sl = slow_list(c);
    const std::vector<uint64_t> key = {sh.wheel ? 1ull : 0ull, sh.tile_log2, sh.wg, sh.med_div, sh.wave_div,
                                       mp.flags, mp.proof_stride, np, tiles, grid, p_small, flags,
                                       (uint64_t)(uintptr_t)c->d_rec, (uint64_t)(uintptr_t)c->d_tinfo,
                                       (uint64_t)(uintptr_t)c->d_wins,
                                       (uint64_t)(uintptr_t)c->d_out, (uint64_t)(uintptr_t)c->d_r0,
                                       (uint64_t)(uintptr_t)c->d_rec_lo, (uint64_t)(uintptr_t)c->d_rec_hi,
                                       (uint64_t)(uintptr_t)c->d_wcopy};
    hipGraphExec_t exec = nullptr;
    for (auto& g : c->graphs)
        if (g.key == key) exec = g.exec;
    if (!exec) {
        if (c->graphs.size() >= GRAPH_CACHE) {
            (void)hipGraphExecDestroy(c->graphs.front().exec);
            c->graphs.erase(c->graphs.begin());
        }
        hipGraph_t g = nullptr;
        HCK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = launch_fill(s, c->d_slow_ctl, 2, 0);
        if (e == hipSuccess) e = launch_fill(s, c->d_counter, 1, 0);
        // the residue launches cover the whole table: a piece's own count (read on the device)
        // bounds the work, so chunks anywhere in the range share the graph
        if (sh.wheel) {
            if (e == hipSuccess) e = launch_residues_wheel(s, T, c->d_wins, np, c->nsp, cls_len, c->d_r0);
            if (e == hipSuccess)
                e = launch_wheel(s, T, mp, c->d_wins, np, tiles, cls_len, (uint32_t)grid, sh.wg, c->d_counter, sl,
                                 c->d_rec, c->d_r0, c->d_tinfo);
        } else {
            if (e == hipSuccess) e = launch_residues(s, T, c->d_wins, np, c->nsp, c->d_r0);
            if (e == hipSuccess)
                e = launch_fused(s, T, mp, c->d_wins, np, tiles, te, (uint32_t)grid, c->d_counter, sl, c->d_rec,
                                 c->d_r0, c->d_wcopy);
        }
        if (e == hipSuccess) e = launch_slow(s, T, mp, sl, c->d_slow_res, c->d_wins, nullptr, nullptr, 0);
        if (e == hipSuccess)
            e = launch_reduce(s, mp, c->d_rec, c->d_rec_lo, c->d_rec_hi, np, sl, c->d_slow_res, c->d_out);
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            HCK(c, e);
        }
        HCK(c, e2);
        hipError_t e3 = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HCK(c, e3);
        c->graphs.push_back(gb_ctx::Graph{key, exec});
    }
    HCK(c, hipEventRecord(c->ev0, s));
    HCK(c, hipGraphLaunch(exec, s));
    HCK(c, hipEventRecord(c->ev1, s));
    HCK(c, hipMemcpyAsync(c->h_out, c->d_out, np * sizeof(StatRec), hipMemcpyDeviceToHost, s));
    HCK(c, hipMemcpyAsync(c->h_ctl, c->d_slow_ctl, 16, hipMemcpyDeviceToHost, s));
    HCK(c, hipStreamSynchronize(s));
    float f = 0;
    HCK(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    if (ms) *ms += f;
    *overflow = reinterpret_cast<const uint32_t*>(c->h_ctl)[2] != 0;
    return GB_OK;
}

// Runs pieces of one shape in launches of at most CHUNK_TILES tiles and LAUNCH_PIECES pieces,
// merging each piece's statistics into out[owner].  A launch whose slow list overflowed runs
// again with its pieces halved (a single tile holds fewer n than the list).  stop_on_fail: stop
// after the first launch with a failure.  *done (nullable) counts the evens finished.
int run_pieces(gb_ctx* c, const Shape& sh, const std::vector<Piece>& pieces, uint64_t p_small, uint32_t flags,
               gb_seg_stats* out, bool stop_on_fail, uint64_t* done, volatile uint64_t* progress, double* ms,
               bool* failed) {
    size_t i = 0;
    while (i < pieces.size()) {
        size_t j = i;
        uint64_t tiles = 0;
        while (j < pieces.size() && j - i < LAUNCH_PIECES) {
            const uint64_t nt = piece_tiles(sh, pieces[j].start, pieces[j].evens);
            if (j > i && tiles + nt > launch_tiles(sh)) break;
            tiles += nt;
            ++j;
        }
        bool overflow = false;
        int rc = run_launch(c, sh, &pieces[i], (uint32_t)(j - i), p_small, flags, ms, &overflow);
        if (rc) return rc;
        if (overflow) {
            std::vector<Piece> halves;
            for (size_t k = i; k < j; ++k) {
                const uint64_t nt = piece_tiles(sh, pieces[k].start, pieces[k].evens);
                if (nt == 1 && j - i == 1)
                    return fail(c, GB_ERANGE, "slow-path list overflow within one tile");
                cut_pieces(sh, pieces[k].start, pieces[k].start + 2 * (pieces[k].evens - 1), pieces[k].owner,
                           std::max<uint64_t>(1, (nt + 1) / 2), halves);
            }
            // one launch per half: each holds at most half the tiles that overflowed
            for (const Piece& h : halves) {
                rc = run_pieces(c, sh, std::vector<Piece>{h}, p_small, flags, out, stop_on_fail, done, progress, ms,
                                failed);
                if (rc || (*failed && stop_on_fail)) return rc;
            }
            i = j;
            continue;
        }
        for (size_t k = i; k < j; ++k) {
            gb_seg_stats one{};
            to_stats(c->h_out[k - i], 0, &one);
            add_stats(&out[pieces[k].owner], one);
            if (one.failures) *failed = true;
            if (done) *done += one.evens;
        }
        if (progress && done) *progress = *done;
        if (*failed && stop_on_fail) return GB_OK;
        i = j;
    }
    return GB_OK;
}

// How gb_verify_range's batched form cuts [start, limit]: (shape, pieces) per kind, ascending.
struct Part {
    Shape sh;
    std::vector<Piece> pieces;
};

int plan_batched(gb_ctx* c, uint64_t start, uint64_t limit, const gb_tuning& t, const gb_tuning* auto_pick,
                 std::vector<Part>& parts) {
    std::vector<std::pair<bool, std::pair<uint64_t, uint64_t>>> kinds;
    split_kinds(start, limit, kinds);
    // the autotuner's pick serves the part it was timed on: the larger one
    size_t dom = 0;
    for (size_t k = 1; k < kinds.size(); ++k)
        if (kinds[k].second.second - kinds[k].second.first > kinds[dom].second.second - kinds[dom].second.first)
            dom = k;
    for (size_t k = 0; k < kinds.size(); ++k) {
        Part P;
        const bool wheel = kinds[k].first;
        if (auto_pick && k == dom)
            P.sh = shape_of(wheel, auto_pick->tile_log2, auto_pick->wg_per_cu, auto_pick->lane_item_div,
                            auto_pick->wave_item_div, t.mr_proof_stride);
        else if (auto_pick) P.sh = shape_of(wheel, 0, 0, 0, 0, t.mr_proof_stride);
        else P.sh = shape_of(wheel, t.tile_log2, t.wg_per_cu, t.lane_item_div, t.wave_item_div, t.mr_proof_stride);
        int rc = check_shape(c, P.sh);
        if (rc) return rc;
        cut_pieces(P.sh, kinds[k].second.first, kinds[k].second.second, 0, launch_tiles(P.sh), P.pieces);
        parts.push_back(std::move(P));
    }
    return GB_OK;
}

// Plain form over [start, limit]: segments of seg evens, stopping at the first failure.  Two
// segments are in flight: segment k + 1 is queued before the host waits for segment k, so the
// device never idles on the host's per-segment read-back (the reference syncs the stream after
// every segment, goldbach.cu:462-463).
int run_plain(gb_ctx* c, uint64_t start, uint64_t limit, uint64_t seg, const MarkParams& mp,
                     gb_seg_stats* tot, volatile uint64_t* progress) {
    const uint64_t first = std::min(seg, (limit - start) / 2 + 1);
    int rc = ensure_plain(c, first, false);
    if (rc) return rc;
    // two segments in flight: on two streams and buffer sets when the second set fits (a range
    // of one segment, or a context that times the kernels apart, keeps one stream)
    SegBufs B[2] = {ctx_bufs(c), ctx_bufs(c)};
    hipStream_t st[2] = {c->stream, c->stream};
    if (first < (limit - start) / 2 + 1 && !(c->cfg.flags & GB_CFG_KERNEL_TIMES) && ensure_plain2(c, first) &&
        (c->plain2.stream || hipStreamCreate(&c->plain2.stream) == hipSuccess)) {
        B[1] = plain2_bufs(c);
        st[1] = c->plain2.stream;
    }
    uint64_t done = 0;
    uint64_t a = start, queued = 0, finished = 0;
    uint64_t cnt_of[2] = {0, 0};
    bool more = true;
    // the device time of the range: the union of its segments' [start, end] clock intervals
    uint64_t busy = 0, end_max = 0;
    const double ms0 = tot->kernel_ms;
    auto queue_next = [&]() -> int {
        const uint64_t left = (limit - a) / 2 + 1;
        const uint64_t cnt = std::min(left, seg);
        const uint32_t slot = (uint32_t)(queued & 1);
        const SegPlan sp = plan_segment(c, a, cnt, c->seg_slot[slot].h_win);
        // the verified map is written, as the reference does
        int r = enqueue_segment(c, st[slot], B[slot], sp, slot, a, cnt, mp, true, false);
        if (r) return r;
        cnt_of[slot] = cnt;
        ++queued;
        if (cnt == left) more = false;
        else a += 2 * cnt;
        return GB_OK;
    };
    rc = queue_next();
    while (rc == GB_OK && finished < queued) {
        if (more) rc = queue_next();
        if (rc) break;
        const uint32_t slot = (uint32_t)(finished & 1);
        StatRec r;
        double ms = 0;
        uint64_t clk[2] = {0, 0};
        rc = finish_segment(c, st[slot], slot, B[slot].sl.cap, &r, &ms, clk);
        if (rc) break;
        ++finished;
        gb_seg_stats one{};
        to_stats(r, ms, &one);
        const uint64_t from = std::max(clk[0], end_max);
        if (clk[1] > from) busy += clk[1] - from;
        end_max = std::max(end_max, clk[1]);
        one.kernel_ms = 0;
        add_stats(tot, one);
        tot->kernel_ms = ms0 + (double)busy / c->wall_khz;
        done += cnt_of[slot];
        if (progress) *progress = done;
        if (one.failures) break;        // the reference stops at the first failure (goldbach.cu:420,476-480)
    }
    // a segment queued past a failure or an error runs out; its results are not counted
    HCK(c, hipStreamSynchronize(st[0]));
    HCK(c, hipStreamSynchronize(st[1]));
    return rc;
}

}  // namespace

extern "C" {


int gb_autotune(gb_ctx* c, uint64_t start, uint64_t limit, uint64_t p_small, uint32_t form, uint32_t flags,
                gb_tuning* best, double* evens_per_s) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!best) return fail(c, GB_EINVAL, "gb_autotune: best is NULL");
    int rc = normalise_range(c, &start, &limit);
    if (rc) return rc;
    if (form != GB_FORM_PLAIN && form != GB_FORM_BATCHED) return fail(c, GB_EINVAL, "unknown form");
    flags &= GB_WANT_PMIN;
    const std::vector<uint64_t> key = {form, start, limit, p_small, flags};
    for (auto& e : c->tune_cache)
        if (e.first == key) {
            *best = e.second;
            if (evens_per_s) *evens_per_s = 0;
            return GB_OK;
        }
    HCK(c, hipSetDevice(c->device));
    // the batched form is timed on the part of the range its larger kernel serves (the wheel up to
    // WHEEL_N_MAX, the fused kernel above), exactly as gb_verify_range will split it
    std::vector<std::pair<bool, std::pair<uint64_t, uint64_t>>> kinds;
    split_kinds(start, limit, kinds);
    bool wheel = kinds[0].first;
    uint64_t a = start, b = limit;
    if (form == GB_FORM_BATCHED) {
        size_t dom = 0;
        for (size_t k = 1; k < kinds.size(); ++k)
            if (kinds[k].second.second - kinds[k].second.first > kinds[dom].second.second - kinds[dom].second.first)
                dom = k;
        wheel = kinds[dom].first;
        a = kinds[dom].second.first;
        b = kinds[dom].second.second;
    }
    const Shape dsh = shape_of(wheel, 0, 0);
    gb_tuning def{};
    def.form = form;
    def.flags = flags;
    def.tile_log2 = dsh.tile_log2;
    def.wg_per_cu = dsh.wg;
    def.seg_evens = std::min<uint64_t>(DEFAULT_SEG, c->cfg.max_seg_evens);
    const uint64_t evens = (b - a) / 2 + 1;
    // four sample windows at 1/8, 3/8, 5/8 and 7/8 of the range: the sieve's cost grows with q
    const uint64_t win = std::min<uint64_t>(evens / 4, form == GB_FORM_BATCHED ? (1ull << 30) : (1ull << 29));
    if (win < (1ull << 22)) {
        *best = def;
        if (evens_per_s) *evens_per_s = 0;
        return GB_OK;
    }
    std::vector<WinDesc> sample;
    for (int k = 0; k < 4; ++k) {
        const uint64_t mid = (evens / 8) * (2 * k + 1);
        const uint64_t first = std::min(mid - std::min(mid, win / 2), evens - win);
        sample.push_back(WinDesc{a + 2 * first, win, 0, 0, 0, 0, 0, 0, 0});
    }
    std::vector<gb_tuning> cand;
    if (form == GB_FORM_BATCHED) {
        for (uint32_t tl = 16; tl <= 20; ++tl) {
            ItemTable it;
            rc = wheel ? build_items(c, wheel_class_len(tl) + WH_HALO + 64, &it, true)
                       : build_items(c, (1u << tl) + HALO + WIN_PAD, &it);
            if (rc) return rc;
            for (uint32_t wg = 1; wg <= 2; ++wg) {
                const size_t lds = wheel ? wheel_lds_bytes(it.n_med, wg) : fused_lds_bytes(1u << tl, it.n_med);
                if (lds * wg > LDS_PER_CU || check_shape(nullptr, Shape{wheel, tl, wg}) != GB_OK) continue;
                gb_tuning t = def;
                t.tile_log2 = tl;
                t.wg_per_cu = wg;
                cand.push_back(t);
            }
        }
    } else {
        for (uint32_t sl = 24; sl <= 29; ++sl) {
            if ((1ull << sl) > c->cfg.max_seg_evens) break;
            gb_tuning t = def;
            t.seg_evens = 1ull << sl;
            cand.push_back(t);
        }
        if (cand.empty()) cand.push_back(def);
    }
    const MarkParams mp = mark_params(c, p_small, flags);
    double best_s = 0;
    gb_tuning pick = def;
    // each candidate runs once to build its schedule and graph, then twice timed; the fastest wins
    auto time_candidates = [&](const std::vector<gb_tuning>& cands) -> int {
        for (const gb_tuning& t : cands) {
            double secs = 1e30;
            for (int rep = 0; rep < 3; ++rep) {
                const auto t0 = std::chrono::steady_clock::now();
                if (form == GB_FORM_BATCHED) {
                    const Shape sh{wheel, t.tile_log2, t.wg_per_cu, t.lane_item_div, t.wave_item_div, t.mr_proof_stride};
                    std::vector<Piece> pcs;
                    for (const WinDesc& w : sample)
                        cut_pieces(sh, w.start, w.start + 2 * (w.evens - 1), 0, launch_tiles(sh), pcs);
                    gb_seg_stats sink{};
                    double ms = 0;
                    bool failed = false;
                    const int r = run_pieces(c, sh, pcs, p_small, flags, &sink, false, nullptr, nullptr, &ms, &failed);
                    if (r) return r;
                } else {
                    for (const WinDesc& w : sample) {
                        gb_seg_stats tot{};
                        const int r = run_plain(c, w.start, w.start + 2 * (w.evens - 1), t.seg_evens, mp, &tot, nullptr);
                        if (r) return r;
                    }
                }
                const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (rep > 0) secs = std::min(secs, dt);
            }
            const double rate = 4.0 * (double)win / secs;
            if (rate > best_s) {
                best_s = rate;
                pick = t;
            }
        }
        return GB_OK;
    };
    rc = time_candidates(cand);
    if (rc) return rc;
    if (form == GB_FORM_BATCHED) {
        // second round: the medium primes' work-item bounds on the fastest tile shape
        const gb_tuning shape = pick;
        const uint32_t tl = shape.tile_log2;
        const uint32_t tb = wheel ? wheel_class_len(tl) + WH_HALO + 64 : (1u << tl) + HALO + WIN_PAD;
        const std::vector<uint32_t> waves = wheel ? std::vector<uint32_t>{WAVE_ITEM_DIV, 2 * WAVE_ITEM_DIV}
                                                  : std::vector<uint32_t>{0u};
        std::vector<gb_tuning> items;
        const uint32_t md0 = wheel ? WHEEL_LANE_ITEM_DIV : LANE_ITEM_DIV;
        for (uint32_t md : {md0 / 2, md0, md0 * 2})
            for (uint32_t wd : waves) {
                if (md == md0 && (wd == WAVE_ITEM_DIV || wd == 0)) continue;   // the first round's
                ItemTable it;
                rc = build_items(c, tb, &it, wheel, md, wd);
                if (rc) return rc;
                const size_t lds = wheel ? wheel_lds_bytes(it.n_med, shape.wg_per_cu) : fused_lds_bytes(1u << tl, it.n_med);
                if (lds * shape.wg_per_cu > LDS_PER_CU) continue;
                gb_tuning t = shape;
                t.lane_item_div = md;
                t.wave_item_div = wd;
                items.push_back(t);
            }
        rc = time_candidates(items);
        if (rc) return rc;
    }
    c->tune_cache.emplace_back(key, pick);
    *best = pick;
    if (evens_per_s) *evens_per_s = best_s;
    return GB_OK;
}

// The tuning a range call runs with: the caller's, or with GB_TUNE_AUTO the autotuner's pick
// (*auto_pick set; it applies to the part of the range it was timed on).
static int resolve_tuning(gb_ctx* c, uint64_t start, uint64_t limit, uint64_t p_small, const gb_tuning* tun,
                          gb_tuning* t, gb_tuning* pick, bool* have_pick) {
    *t = tun ? *tun : gb_tuning{};
    *have_pick = false;
    if (t->form != GB_FORM_PLAIN && t->form != GB_FORM_BATCHED) return fail(c, GB_EINVAL, "unknown form");
    if (t->flags & GB_TUNE_AUTO) {
        int rc = gb_autotune(c, start, limit, p_small, t->form, t->flags, pick, nullptr);
        if (rc) return rc;
        *have_pick = true;
        t->seg_evens = pick->seg_evens;
    }
    return GB_OK;
}

int gb_range_plan(gb_ctx* c, uint64_t start, uint64_t limit, uint64_t p_small, const gb_tuning* tun, gb_chunk* out,
                  uint64_t cap, uint64_t* count) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!count) return fail(c, GB_EINVAL, "gb_range_plan: count is NULL");
    *count = 0;
    int rc = normalise_range(c, &start, &limit);
    if (rc) return rc;
    HCK(c, hipSetDevice(c->device));
    gb_tuning t{}, pick{};
    bool have_pick = false;
    rc = resolve_tuning(c, start, limit, p_small, tun, &t, &pick, &have_pick);
    if (rc) return rc;
    uint64_t n = 0;
    auto emit = [&](const gb_chunk& ch) {
        if (out && n < cap) out[n] = ch;
        ++n;
    };
    if (t.form == GB_FORM_PLAIN) {
        const uint64_t seg = t.seg_evens ? t.seg_evens : std::min<uint64_t>(DEFAULT_SEG, c->cfg.max_seg_evens);
        if (seg > c->cfg.max_seg_evens) return fail(c, GB_ERANGE, "seg_evens above max_seg_evens");
        for (uint64_t a = start;;) {
            const uint64_t left = (limit - a) / 2 + 1, cnt = std::min(left, seg);
            emit(gb_chunk{a, a + 2 * (cnt - 1), 0, GB_KERNEL_PLAIN, 0, 0, 0});
            if (cnt == left) break;
            a += 2 * cnt;
        }
    } else {
        std::vector<Part> parts;
        rc = plan_batched(c, start, limit, t, have_pick ? &pick : nullptr, parts);
        if (rc) return rc;
        for (const Part& P : parts)
            for (const Piece& pc : P.pieces)
                emit(gb_chunk{pc.start, pc.start + 2 * (pc.evens - 1), piece_tiles(P.sh, pc.start, pc.evens),
                              P.sh.wheel ? GB_KERNEL_WHEEL : GB_KERNEL_FUSED, P.sh.tile_log2, P.sh.wg, 0});
    }
    *count = n;
    return GB_OK;
}

int gb_verify_range(gb_ctx* c, uint64_t start, uint64_t limit, uint64_t p_small, const gb_tuning* tun,
                    gb_seg_stats* st, volatile uint64_t* progress) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    int rc = normalise_range(c, &start, &limit);
    if (rc) return rc;
    HCK(c, hipSetDevice(c->device));
    gb_tuning t{}, pick{};
    bool have_pick = false;
    rc = resolve_tuning(c, start, limit, p_small, tun, &t, &pick, &have_pick);
    if (rc) return rc;
    gb_seg_stats tot{};
    if (t.form == GB_FORM_BATCHED) {
        std::vector<Part> parts;
        rc = plan_batched(c, start, limit, t, have_pick ? &pick : nullptr, parts);
        if (rc) return rc;
        uint64_t done = 0;
        double ms = 0;
        bool failed = false;
        for (const Part& P : parts) {
            rc = run_pieces(c, P.sh, P.pieces, p_small, t.flags & GB_WANT_PMIN, &tot, true, &done, progress, &ms,
                            &failed);
            if (rc) return rc;
            if (failed) break;   // the reference stops at its first failure (goldbach.cu:420,476-480)
        }
        tot.kernel_ms = ms;
        if (st) *st = tot;
        return GB_OK;
    }
    uint64_t seg = t.seg_evens ? t.seg_evens : std::min<uint64_t>(DEFAULT_SEG, c->cfg.max_seg_evens);
    if (seg > c->cfg.max_seg_evens) return fail(c, GB_ERANGE, "seg_evens above max_seg_evens");
    const MarkParams mp = mark_params(c, p_small, t.flags);
    rc = run_plain(c, start, limit, seg, mp, &tot, progress);
    if (rc) return rc;
    if (st) *st = tot;
    return GB_OK;
}

int gb_verify_windows(gb_ctx* c, const gb_window* windows, uint32_t count, uint64_t p_small, uint32_t flags,
                      gb_seg_stats* per_window, gb_seg_stats* total) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!windows || count == 0) return fail(c, GB_EINVAL, "no windows");
    for (uint32_t i = 0; i < count; ++i) {
        const uint64_t a = windows[i].start, e = windows[i].evens;
        if (a < 4 || (a & 1) || e == 0 || e - 1 > (~0ull - 1 - a) / 2)
            return fail(c, GB_EINVAL, "window " + std::to_string(i) + " outside [4, 2^64-2] or not even");
    }
    HCK(c, hipSetDevice(c->device));
    // each window is split where the wheel stops and cut on its kernel's tile grid; the pieces of
    // one kind share launches
    const Shape shw = shape_of(true, 0, 0), shf = shape_of(false, 0, 0);
    std::vector<Piece> pw, pf;
    for (uint32_t i = 0; i < count; ++i) {
        std::vector<std::pair<bool, std::pair<uint64_t, uint64_t>>> kinds;
        const uint64_t a = windows[i].start;
        split_kinds(a, a + 2 * (windows[i].evens - 1), kinds);
        for (auto& k : kinds)
            cut_pieces(k.first ? shw : shf, k.second.first, k.second.second, i, launch_tiles(k.first ? shw : shf),
                       k.first ? pw : pf);
    }
    std::vector<gb_seg_stats> per(count);
    double ms = 0;
    bool failed = false;
    int rc = run_pieces(c, shw, pw, p_small, flags & GB_WANT_PMIN, per.data(), false, nullptr, nullptr, &ms, &failed);
    if (rc) return rc;
    rc = run_pieces(c, shf, pf, p_small, flags & GB_WANT_PMIN, per.data(), false, nullptr, nullptr, &ms, &failed);
    if (rc) return rc;
    gb_seg_stats tot{};
    for (uint32_t i = 0; i < count; ++i) {
        if (per_window) per_window[i] = per[i];
        add_stats(&tot, per[i]);
    }
    tot.kernel_ms = ms;
    if (total) *total = tot;
    return GB_OK;
}

int gb_count_partitions(gb_ctx* c, uint64_t start, uint64_t count, uint64_t* counts_out) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!counts_out) return fail(c, GB_EINVAL, "counts_out is NULL");
    if (count == 0) return GB_OK;
    if (start < 4 || (start & 1)) return fail(c, GB_EINVAL, "start must be even and >= 4");
    if (start > COUNT_N_MAX || count - 1 > (COUNT_N_MAX - start) / 2) return fail(c, GB_ERANGE, "partition counts reach n <= 2^36 only");
    HCK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const uint64_t n_end = start + 2 * (count - 1);
    std::fill(counts_out, counts_out + count, 0ull);
    if (n_end < 6) {                       // only n = 4: 2 + 2
        counts_out[0] = 1;
        return GB_OK;
    }
    // P: odd primes q in [3, n_end - 3] (bit a <-> 2a + 3), M = (n_end - 6) / 2 its last index
    const uint64_t q_high = n_end - 3, M = (q_high - 3) / 2;
    const uint64_t nbits = M + 1, pwords = (nbits + 63) / 64 + 2;
    ItemTable it;
    int rc = build_items(c, PLAIN_TILE_BITS, &it);
    if (rc) return rc;
    const Tables T = make_tables(c, it);
    uint64_t *d_P = nullptr, *d_R = nullptr, *d_cnt = nullptr;
    auto release = [&]() {
        dfree(c, d_P);
        dfree(c, d_R);
        dfree(c, d_cnt);
    };
    if (dalloc(c, &d_P, pwords) != hipSuccess || dalloc(c, &d_R, pwords) != hipSuccess ||
        dalloc(c, &d_cnt, count) != hipSuccess) {
        release();
        return hip_fail(c, hipErrorOutOfMemory, __FILE__, __LINE__);
    }
    rc = ensure_wins(c, 1);
    if (rc == GB_OK) {
        WinDesc wd{0, 0, 0, 0, 0, 0, 0, 0, 0};
        set_residues(c, &wd, 3, q_high, 0);
        rc = ensure_r0(c, wd.np);
        if (rc == GB_OK) {
            hipError_t e = launch_fill(s, d_P, pwords, 0);
            if (e == hipSuccess) e = launch_fill(s, d_cnt, count, 0);
            if (e == hipSuccess) e = hipMemcpyAsync(c->d_wins, &wd, sizeof wd, hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = launch_residues(s, T, c->d_wins, 1, wd.np, c->d_r0);
            if (e == hipSuccess) e = launch_sieve_window(s, T, 3, 0, nbits, PLAIN_TILE_BITS, d_P, 0, c->d_r0);
            if (e == hipSuccess) e = hipEventRecord(c->ev0, s);
            if (e == hipSuccess) e = launch_count_partitions(s, d_P, d_R, M, pwords, start, n_end, d_cnt);
            if (e == hipSuccess) e = hipEventRecord(c->ev1, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(counts_out, d_cnt, count * 8, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = fail(c, GB_EDEVICE, std::string("partition counts: ") + hipGetErrorString(e));
        }
    }
    release();
    if (rc) return rc;
    if (start == 4) counts_out[0] += 1;   // 4 = 2 + 2, the only partition with p = 2
    return GB_OK;
}

int gb_min_partition_batch(gb_ctx* c, const uint64_t* n, uint32_t count, uint64_t* p_out) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    return gbh::min_partition_bounded(c, n, count, c->cfg.phase2_limit, p_out);
}

}  // extern "C"

namespace gbh {

int min_partition_bounded(gb_ctx* c, const uint64_t* n, uint32_t count, uint64_t bound, uint64_t* p_out) {
    if (count == 0) return GB_OK;
    if (!n || !p_out) return fail(c, GB_EINVAL, "NULL buffer");
    HCK(c, hipSetDevice(c->device));
    if (count > c->q_cap) {
        dfree(c, c->d_qn);
        dfree(c, c->d_qp);
        c->d_qn = c->d_qp = nullptr;
        c->q_cap = 0;
        HCK(c, dalloc(c, &c->d_qn, count));
        HCK(c, dalloc(c, &c->d_qp, count));
        c->q_cap = count;
    }
    ItemTable it;
    int rc = build_items(c, PLAIN_TILE_BITS, &it);
    if (rc) return rc;
    const Tables T = make_tables(c, it);
    HCK(c, hipMemcpyAsync(c->d_qn, n, count * 8ull, hipMemcpyHostToDevice, c->stream));
    HCK(c, launch_min_partition(c->stream, T, c->d_qn, count, bound, c->d_qp));
    HCK(c, hipMemcpyAsync(p_out, c->d_qp, count * 8ull, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    return GB_OK;
}

}  // namespace gbh
