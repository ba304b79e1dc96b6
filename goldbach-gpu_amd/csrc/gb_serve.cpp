// gb_serve.cpp -- the continuous-batching serving engine of goldbach_gpu.h.
//
// A server's slots form two lanes, each with its own engine thread and HIP stream.  A slot's query
// state (n and scan position, 16 bytes) lives in device memory, allocated from the
// context's budget, so the number of resident queries is what the device memory granted
// to the server holds.  Per slot the host keeps two pinned words the step kernel reads
// and writes directly: the n of a query admitted this step, and the step's answer.
// Each iteration (a "step") a lane's engine
//   1. moves queued queries into its free slots (the lowest free ones first),
//   2. launches serve_step_kernel over its occupied range, new and in-flight queries
//      alike (one wave per slot, up to chunks_per_step chunks of 64 candidate primes
//      from the query's saved position, which stays on the device),
//   3. retires the queries that finished; the others keep their slot.
// The two lanes' host work runs on two cores and their steps overlap on the device.
// Submitters and pollers only touch host queues under a mutex, so arrivals overlap
// with the step running on the device and join at the next step.  Queues are vectors
// appended and drained in bulk, and the engine builds a step's results without the
// lock, so a burst costs a few nanoseconds of host work per query.
// An idle engine polls an arrival counter for up to 200 us before it sleeps on
// its condition variable, so a query arriving at an idle server does not wait for
// the thread to be woken (the price: one host core busy while queries keep coming).
//
// Replaces single_check's flow (single_check.cu:136-319 of the reference:
// CPU sieve of p chunks, one launch per chunk, first finder wins) with a
// resident-table request loop whose answer is the minimal p.
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gb_ctx.hpp"

using namespace gbk;
using namespace gbh;

namespace {

constexpr uint32_t DEFAULT_CHUNKS = 4;
constexpr uint64_t DEFAULT_QUEUE = 1ull << 22;
constexpr uint64_t IDLE_SPIN_NS = 200000;   // an idle engine polls this long for arrivals before it sleeps
constexpr uint64_t SLOT_DEVICE_BYTES = sizeof(ServeSlot);   // device memory one resident query holds

uint64_t now_ns() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

struct Query {
    uint64_t id;
    uint64_t n;
    uint64_t submit_ns;
};

struct SlotInfo {   // host side of an occupied slot (n = 0: free)
    uint64_t id;
    uint64_t n;
    uint64_t submit_ns;
    uint32_t steps;
};

// A FIFO of trivially copyable records appended and drained in bulk.
template <typename T>
struct Fifo {
    std::vector<T> v;
    size_t head = 0;
    size_t size() const { return v.size() - head; }
    bool empty() const { return head == v.size(); }
    // Drops what was drained: for free once all of it was, otherwise only when the part still
    // queued is a sixteenth of the part drained, so the move it costs stays amortised.
    void compact() {
        if (head == v.size()) {
            v.clear();
            head = 0;
        } else if (head > 65536 && v.size() - head < head / 16) {
            v.erase(v.begin(), v.begin() + (ptrdiff_t)head);
            head = 0;
        }
    }
};

// Free slots as a bitmap; take() hands out the lowest, so the occupied slots stay packed at the
// bottom of the table and a step launches over no more slots than it must.
struct FreeSlots {
    std::vector<uint64_t> bits;
    size_t hint = 0;   // no free slot below word `hint`
    uint32_t count;
    explicit FreeSlots(uint32_t n) : bits((n + 63) / 64, ~0ull), count(n) {
        if (n % 64) bits.back() = (1ull << (n % 64)) - 1;
    }
    size_t size() const { return count; }
    uint32_t take() {   // callers check size() first
        while (!bits[hint]) ++hint;
        const uint32_t b = (uint32_t)__builtin_ctzll(bits[hint]);
        bits[hint] &= bits[hint] - 1;
        --count;
        return (uint32_t)(hint * 64 + b);
    }
    void give(uint32_t i) {
        bits[i / 64] |= 1ull << (i % 64);
        hint = std::min<size_t>(hint, i / 64);
        ++count;
    }
};

// A lane: half of the server's slots with its own engine thread and stream.  Its query
// state is in device memory; per slot two pinned words the step kernel reads and writes:
// the n of a query admitted this step, and the step's answer.
struct Lane {
    uint32_t size = 0;              // slots
    ServeSlot* d_slots = nullptr;   // within the server's device table
    uint64_t* adm = nullptr;
    ServeDone* done = nullptr;
    hipStream_t stream = nullptr;
    std::thread engine;
};

}  // namespace

struct gb_server {
    gb_ctx* ctx = nullptr;
    int device = 0;
    Tables T{};
    uint64_t bound = 0;
    uint32_t slots = 0;
    uint32_t chunks = 0;
    uint64_t queue_cap = 0;
    ServeSlot* d_slots = nullptr;   // query state, device memory (counted against the context's budget)
    Lane lanes[2];                  // two engines, each on half of the slots
    uint64_t lane_resident[2] = {0, 0};
    std::mutex submit_mu;           // one submission at a time, so its ids stay consecutive
    std::mutex mu;
    std::condition_variable cv_work;   // lane 0 waits for queries or stop
    std::condition_variable cv_help;   // lane 1 waits for what lane 0 leaves queued
    std::condition_variable cv_done;   // pollers wait for results
    Fifo<Query> queue;
    Fifo<gb_result> done;              // finished, not yet taken by a poller
    std::mutex poll_mu;                // pollers, one at a time
    Fifo<gb_result> polled;            // taken by pollers, not yet returned
    uint64_t next_id = 0;
    std::atomic<uint64_t> arrivals{0};   // submissions so far: read without the lock by an idle engine
    std::atomic<bool> stopping{false};
    bool stop = false;
    bool failed = false;
    std::string err;
    gb_serve_stats st{};
};

namespace {

void engine_fail(gb_server* s, const std::string& msg) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->failed = true;
    s->err = msg;
    s->cv_done.notify_all();
}

// One lane's engine: retire, admit, launch, wait, on the lane's own slots and stream.
void engine_main(gb_server* s, uint32_t lane) {
    if (hipSetDevice(s->device) != hipSuccess) {
        engine_fail(s, "serving engine: hipSetDevice failed");
        return;
    }
    Lane& L = s->lanes[lane];
    std::vector<SlotInfo> info(L.size, SlotInfo{0, 0, 0, 0});
    FreeSlots free_slots(L.size);
    std::vector<Query> admit;
    std::vector<uint32_t> admitted;
    std::vector<gb_result> results;
    uint32_t hi = 0, resident = 0;   // the lane's slots [0, hi) hold every query resident in it
    uint64_t seen = 0;               // submissions counted at the last admission
    for (;;) {
        if (resident == 0 && lane == 0) {
            // idle: poll for arrivals for a short while, so a query that arrives now does not
            // wait for the thread to be woken; then sleep on the condition variable
            const uint64_t t0 = now_ns();
            while (s->arrivals.load(std::memory_order_acquire) == seen && !s->stopping.load(std::memory_order_acquire) &&
                   now_ns() - t0 < IDLE_SPIN_NS) {
            }
        }
        admit.clear();
        {
            // lane 0 serves arrivals; lane 1 sleeps until lane 0 leaves queries queued (a burst)
            std::unique_lock<std::mutex> lk(s->mu);
            (lane == 0 ? s->cv_work : s->cv_help).wait(lk, [&] { return s->stop || !s->queue.empty() || resident > 0; });
            seen = s->arrivals.load(std::memory_order_relaxed);
            if (s->stop) return;
            const size_t take = std::min(s->queue.size(), free_slots.size());
            admit.assign(s->queue.v.begin() + (ptrdiff_t)s->queue.head,
                         s->queue.v.begin() + (ptrdiff_t)(s->queue.head + take));
            s->queue.head += take;
            s->queue.compact();
            s->st.queued = s->queue.size();
            s->lane_resident[lane] = resident + take;
            s->st.resident = s->lane_resident[0] + s->lane_resident[1];
            s->st.max_resident = std::max<uint64_t>(s->st.max_resident, s->st.resident);
            if (lane == 0 && !s->queue.empty()) s->cv_help.notify_one();   // the other lane has room
        }
        admitted.clear();
        for (const Query& q : admit) {
            const uint32_t i = free_slots.take();
            info[i] = SlotInfo{q.id, q.n, q.submit_ns, 0};
            L.adm[i] = q.n;
            admitted.push_back(i);
            hi = std::max(hi, i + 1);
        }
        resident += (uint32_t)admit.size();
        if (resident == 0) continue;
        const uint32_t m = hi;
        hipError_t e = launch_serve_step(L.stream, s->T, L.adm, L.d_slots, m, s->bound, s->chunks, L.done);
        if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
        if (e != hipSuccess) {
            engine_fail(s, std::string("serving engine: ") + hipGetErrorString(e));
            return;
        }
        const uint64_t t = now_ns();
        for (uint32_t i : admitted) L.adm[i] = 0;
        results.clear();
        for (uint32_t i = 0; i < m; ++i) {
            SlotInfo& q = info[i];
            if (!q.n) continue;
            q.steps += 1;
            const ServeDone d = L.done[i];
            if (d.state != 1) continue;
            results.push_back(gb_result{q.id, q.n, d.p, q.submit_ns, t, q.steps, 0});
            q.n = 0;
            free_slots.give(i);
        }
        const uint64_t finished = results.size();
        const uint64_t in_step = resident;
        resident -= (uint32_t)finished;
        while (hi > 0 && info[hi - 1].n == 0) --hi;
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->done.empty()) {
            s->done.v.swap(results);   // hand the results over without copying them
            s->done.head = 0;
        } else {
            s->done.v.insert(s->done.v.end(), results.begin(), results.end());
        }
        s->st.steps += 1;
        s->st.step_queries += in_step;
        s->st.completed += finished;
        s->lane_resident[lane] = resident;
        s->st.resident = s->lane_resident[0] + s->lane_resident[1];
        if (finished) s->cv_done.notify_all();
    }
}

void free_server(gb_server* s) {
    for (Lane& L : s->lanes)
        if (L.stream) (void)hipStreamSynchronize(L.stream);
    for (Lane& L : s->lanes) {
        if (L.adm) (void)hipHostFree(L.adm);
        if (L.done) (void)hipHostFree(L.done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    if (s->d_slots) dfree(s->ctx, s->d_slots);
    delete s;
}

}  // namespace

extern "C" {

int gb_serve_open(gb_ctx* c, const gb_serve_config* cfg, gb_server** out) {
    if (!c) return GB_EINVAL;
    if (!out) return fail(c, GB_EINVAL, "gb_serve_open: out is NULL");
    *out = nullptr;
    c->err.clear();
    gb_serve_config k{};
    if (cfg) k = *cfg;
    uint64_t slots = k.slots ? k.slots : c->cfg.serve_slots;
    if (k.mem_budget) slots = std::min<uint64_t>(slots, k.mem_budget / SLOT_DEVICE_BYTES);
    if (slots == 0) return fail(c, GB_ENOMEM, "serving: the device memory budget holds no slot");
    if (slots > (1u << 24)) return fail(c, GB_EINVAL, "serving: more than 2^24 slots");
    HCK(c, hipSetDevice(c->device));
    ItemTable it;
    int rc = build_items(c, PLAIN_TILE_BITS, &it);
    if (rc) return rc;

    gb_server* s = new gb_server();
    s->ctx = c;
    s->device = c->device;
    s->T = make_tables(c, it);
    s->bound = c->cfg.phase2_limit;
    s->slots = (uint32_t)slots;
    s->chunks = k.chunks_per_step ? k.chunks_per_step : DEFAULT_CHUNKS;
    s->queue_cap = k.queue_cap ? k.queue_cap : DEFAULT_QUEUE;
    s->st.slots = s->slots;
    hipError_t e = dalloc(c, &s->d_slots, slots);   // the context's budget counts the slot state
    const uint32_t half = (uint32_t)((slots + 1) / 2);
    for (uint32_t l = 0; l < 2; ++l) {
        Lane& L = s->lanes[l];
        L.size = l == 0 ? half : (uint32_t)slots - half;
        const size_t n = std::max<uint32_t>(L.size, 1);
        if (e == hipSuccess) e = hipStreamCreate(&L.stream);
        if (e == hipSuccess) L.d_slots = s->d_slots + (l == 0 ? 0 : half);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L.adm), n * sizeof(uint64_t), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&L.done), n * sizeof(ServeDone), hipHostMallocDefault);
        if (e == hipSuccess) std::memset(L.adm, 0, n * sizeof(uint64_t));
    }
    if (e == hipSuccess) e = hipMemsetAsync(s->d_slots, 0, slots * sizeof(ServeSlot), s->lanes[0].stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->lanes[0].stream);
    if (e != hipSuccess) {
        const int rc = hip_fail(c, e, __FILE__, __LINE__);
        free_server(s);
        c->err = "serving: set-up failed -- " + c->err;
        return rc;
    }
    try {
        for (uint32_t l = 0; l < 2; ++l)
            if (s->lanes[l].size) s->lanes[l].engine = std::thread(engine_main, s, l);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->stopping.store(true);
        }
        s->cv_work.notify_all();
        s->cv_help.notify_all();
        for (Lane& L : s->lanes)
            if (L.engine.joinable()) L.engine.join();
        free_server(s);
        return fail(c, GB_EDEVICE, "serving: could not start the engine threads");
    }
    c->servers.fetch_add(1);
    *out = s;
    return GB_OK;
}

int gb_serve_submit(gb_server* s, const uint64_t* n, uint32_t count, uint64_t* first_id) {
    if (!s) return GB_EINVAL;
    if (count && !n) return GB_EINVAL;
    for (uint32_t i = 0; i < count; ++i)
        if (n[i] < 4 || (n[i] & 1)) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = "gb_serve_submit: n[" + std::to_string(i) + "] is odd or below 4";
            return GB_EINVAL;
        }
    const uint64_t t = now_ns();
    std::lock_guard<std::mutex> sk(s->submit_mu);   // ids of one submission stay consecutive
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->failed) return GB_EDEVICE;
        // only the engine takes from the queue meanwhile: room now is room for every chunk below
        if (s->queue.size() + count > s->queue_cap) {
            s->err = "gb_serve_submit: queue full";
            return GB_EAGAIN;
        }
        if (first_id) *first_id = s->next_id;
        s->queue.v.reserve(s->queue.v.size() + count);   // one reallocation, not one per growth step
    }
    // appended in chunks, so the engine starts on the first while the rest is copied
    constexpr uint32_t CHUNK = 1u << 15;
    for (uint32_t c0 = 0; c0 < count; c0 += CHUNK) {
        const uint32_t c = std::min(CHUNK, count - c0);
        std::lock_guard<std::mutex> lk(s->mu);
        const size_t at = s->queue.v.size();
        s->queue.v.resize(at + c);
        Query* q = s->queue.v.data() + at;
        for (uint32_t i = 0; i < c; ++i) q[i] = Query{s->next_id + i, n[c0 + i], t};
        s->next_id += c;
        s->arrivals.fetch_add(c, std::memory_order_release);
        s->st.submitted += c;
        s->st.queued = s->queue.size();
        s->cv_work.notify_one();
    }
    return GB_OK;
}

int gb_serve_poll(gb_server* s, gb_result* out, uint32_t max, uint32_t* got, uint32_t wait_ms) {
    if (!s) return GB_EINVAL;
    if (got) *got = 0;
    if (max && !out) return GB_EINVAL;
    // pollers take the engines' finished results whole (a swap under the engines' lock) and copy
    // them out under their own lock, so an engine never waits for a poller's copy
    std::lock_guard<std::mutex> pk(s->poll_mu);
    Fifo<gb_result>& mine = s->polled;
    bool failed = false;
    if (mine.empty()) {
        std::unique_lock<std::mutex> lk(s->mu);
        if (wait_ms && s->done.empty() && !s->failed)
            s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty() || s->failed; });
        mine.v.clear();
        mine.head = 0;
        mine.v.swap(s->done.v);
        std::swap(mine.head, s->done.head);
        failed = s->failed;
    }
    const uint32_t k = (uint32_t)std::min<size_t>(max, mine.size());
    if (k) std::memcpy(out, mine.v.data() + mine.head, k * sizeof(gb_result));
    mine.head += k;
    if (got) *got = k;
    if (failed && k == 0) return GB_EDEVICE;
    return GB_OK;
}

int gb_serve_stats_get(gb_server* s, gb_serve_stats* st) {
    if (!s || !st) return GB_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->st;
    return GB_OK;
}

int gb_serve_close(gb_server* s) {
    if (!s) return GB_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->stopping.store(true, std::memory_order_release);
    }
    s->cv_work.notify_all();
    s->cv_help.notify_all();
    for (Lane& L : s->lanes)
        if (L.engine.joinable()) L.engine.join();
    gb_ctx* c = s->ctx;
    (void)hipSetDevice(s->device);
    free_server(s);
    c->servers.fetch_sub(1);
    return GB_OK;
}

const char* gb_serve_last_error(const gb_server* s) {
    if (!s) return "";
    return s->err.c_str();
}

}  // extern "C"
