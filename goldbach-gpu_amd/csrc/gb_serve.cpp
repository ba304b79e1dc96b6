// gb_serve.cpp -- the continuous-batching serving engine of goldbach_gpu.h.
//
// A server's slots form four lanes, each with its own HIP stream; two engine threads drive
// two lanes each.  A slot's query state (n and scan position, 16 bytes) lives in device
// memory, allocated from the context's budget, so the number of resident queries is what the
// device memory granted to the server holds.  Per slot the host keeps two pinned words the
// step kernel reads and writes directly: the n of a query admitted this step, and the
// step's answer.  An engine takes its lanes in turn; for each it
//   1. waits for the lane's step in flight, if any, and retires the queries that finished
//      (the others keep their slot),
//   2. moves queued queries into the lane's free slots (the lowest free ones first),
//   3. launches serve_step_kernel over the lane's occupied range, new and in-flight queries
//      alike (one wave per slot, up to chunks_per_step chunks of 64 candidate primes
//      from the query's saved position, which stays on the device).
// So an engine's host work on one lane overlaps the step of its other lane on the device,
// the two engines' host work runs on two cores, and their steps overlap on the device.
// Submitters and pollers only touch host queues under a mutex, so arrivals overlap
// with the step running on the device and join at the next step.  Queues are vectors
// appended and drained in bulk, and the engine builds a step's results without the
// lock and hands them over as one block, so a burst costs a few nanoseconds of host work
// per query.
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
#include <deque>
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

// A lane: a quarter of the server's slots, with its host-side state.  Its query state is in
// device memory; per slot two pinned words the step kernel reads and writes: the n of a
// query admitted this step, and the step's answer.
struct Lane {
    uint32_t size = 0;              // slots
    ServeSlot* d_slots = nullptr;   // within the server's device table
    uint64_t* adm = nullptr;
    ServeDone* done = nullptr;
    hipStream_t stream = nullptr;   // the lane's own: a step admitted while the other lane's
                                    // runs does not queue behind it
    uint32_t* d_blocks = nullptr;   // the step's workgroups that have finished (device memory)
    uint64_t* h_seq = nullptr;      // pinned: the last step that has finished (written by the step)
    uint64_t seq = 0;               // the lane's steps launched
    std::vector<SlotInfo> info;
    FreeSlots free_slots{0};
    std::vector<uint32_t> admitted;   // slots whose adm word the step in flight reads
    uint32_t hi = 0;                  // slots [0, hi) hold every query resident in the lane
    uint32_t resident = 0;
    bool in_flight = false;
    uint64_t launched_ns = 0;
};

// An engine: one host thread driving two lanes in turn, so the thread retires and refills
// one lane while the other's step runs on the device.
struct Engine {
    Lane lanes[2];
    std::thread thread;
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
    uint32_t* d_blocks = nullptr;   // the lanes' step completion counters (device memory)
    Engine engines[2];              // two engines, each on half of the slots (two lanes each)
    uint64_t engine_resident[2] = {0, 0};
    std::mutex submit_mu;           // one submission at a time, so its ids stay consecutive
    std::mutex mu;
    std::condition_variable cv_work;   // engine 0 waits for queries or stop
    std::condition_variable cv_help;   // engine 1 waits for what engine 0 leaves queued
    std::condition_variable cv_done;   // pollers wait for results
    Fifo<Query> queue;
    // Finished results travel as whole blocks (one per step that finished any), so handing a
    // step's results to the pollers and taking them are O(1) under the lock whatever the
    // backlog; drained blocks go back to a pool and keep their capacity for later steps.
    std::deque<std::vector<gb_result>> done;   // finished, not yet taken by a poller
    std::vector<std::vector<gb_result>> pool;  // drained blocks for reuse
    size_t pool_max = 0;                       // blocks the pool keeps (about 64 MiB of results)
    std::mutex poll_mu;                        // pollers, one at a time
    std::vector<gb_result> polled;             // the block a poller is returning from
    size_t polled_head = 0;
    uint64_t next_id = 0;
    std::atomic<uint64_t> arrivals{0};   // submissions so far: read without the lock by an idle engine
    std::atomic<uint64_t> queued{0};     // queue length, kept with the queue (read without the lock)
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

// Retires the finished queries of lane L's step (the host has seen its event) and hands
// their results over.
void retire(gb_server* s, uint32_t eng, Lane& L, std::vector<gb_result>& results, uint64_t t_start) {
    const uint64_t t = now_ns();
    for (uint32_t i : L.admitted) L.adm[i] = 0;
    L.admitted.clear();
    results.clear();
    const uint32_t m = L.hi;
    results.reserve(m);
    for (uint32_t i = 0; i < m; ++i) {
        SlotInfo& q = L.info[i];
        if (!q.n) continue;
        q.steps += 1;
        const ServeDone d = L.done[i];
        if (d.state != 1) continue;
        results.push_back(gb_result{q.id, q.n, d.p, q.submit_ns, t, q.steps, 0});
        q.n = 0;
        L.free_slots.give(i);
    }
    const uint64_t finished = results.size();
    const uint64_t in_step = L.resident;
    L.resident -= (uint32_t)finished;
    while (L.hi > 0 && L.info[L.hi - 1].n == 0) --L.hi;
    const uint64_t t_r = now_ns();
    std::lock_guard<std::mutex> lk(s->mu);
    if (finished) {
        s->done.push_back(std::move(results));   // hand the block over; take a drained one back
        results = std::vector<gb_result>();
        if (!s->pool.empty()) {
            results.swap(s->pool.back());
            s->pool.pop_back();
        }
    }
    s->st.queue_ns += now_ns() - t_r;
    s->st.device_ns += t - t_start;
    s->st.retire_ns += t_r - t;
    s->st.steps += 1;
    s->st.step_queries += in_step;
    s->st.completed += finished;
    s->engine_resident[eng] -= finished;
    s->st.resident = s->engine_resident[0] + s->engine_resident[1];
    if (finished) s->cv_done.notify_all();
}

// One engine: for each of its two lanes in turn, retire the step in flight (waiting for it),
// admit queued queries into the lane's free slots and launch its next step.  While one lane's
// step runs, the thread works on the other lane.
void engine_main(gb_server* s, uint32_t eng) {
    if (hipSetDevice(s->device) != hipSuccess) {
        engine_fail(s, "serving engine: hipSetDevice failed");
        return;
    }
    Engine& E = s->engines[eng];
    std::vector<Query> admit;
    std::vector<gb_result> results;
    uint64_t seen = 0;   // submissions counted at the last admission
    for (uint32_t turn = 0;; turn ^= 1) {
        Lane& L = E.lanes[turn];
        Lane& other = E.lanes[turn ^ 1];
        if (L.size == 0) continue;
        if (L.in_flight) {
            // spin on the step's pinned completion word; now and then ask the stream, so a device
            // error ends the wait instead of leaving it spinning
            const volatile uint64_t* done_seq = L.h_seq;
            for (uint32_t spin = 1; *done_seq != L.seq; ++spin) {
                if (spin % 512) continue;
                const hipError_t e = hipStreamQuery(L.stream);
                if (e == hipErrorNotReady) continue;
                if (e != hipSuccess || *done_seq != L.seq) {
                    engine_fail(s, e != hipSuccess ? std::string("serving engine: ") + hipGetErrorString(e)
                                                   : std::string("serving engine: a step ended without its signal"));
                    return;
                }
            }
            // the answers were written before the sequence word: none of their reads may move above
            // the loop that saw it
            std::atomic_thread_fence(std::memory_order_acquire);
            L.in_flight = false;
            retire(s, eng, L, results, L.launched_ns);
        }
        if (L.resident == 0 && other.in_flight && s->queued.load(std::memory_order_acquire) == 0)
            continue;   // nothing for this lane: go back to the step in flight
        const bool busy = L.resident > 0 || other.in_flight;
        if (!busy && eng == 0) {
            // idle: poll for arrivals for a short while, so a query that arrives now does not
            // wait for the thread to be woken; then sleep on the condition variable
            const uint64_t t0 = now_ns();
            while (s->arrivals.load(std::memory_order_acquire) == seen && !s->stopping.load(std::memory_order_acquire) &&
                   now_ns() - t0 < IDLE_SPIN_NS) {
            }
        }
        admit.clear();
        uint64_t t_q0 = now_ns(), t_q1;
        {
            // engine 0 serves arrivals; engine 1 sleeps until engine 0 leaves queries queued (a
            // burst).  An engine with a step in flight does not sleep: it has that step to retire.
            std::unique_lock<std::mutex> lk(s->mu);
            auto ready = [&] { return s->stop || !s->queue.empty() || L.resident > 0 || other.in_flight; };
            if (!ready()) {
                (eng == 0 ? s->cv_work : s->cv_help).wait(lk, ready);
                t_q0 = now_ns();   // the wait for work is idle time, not queue time
            }
            seen = s->arrivals.load(std::memory_order_relaxed);
            if (s->stop) break;
            const size_t take = std::min(s->queue.size(), L.free_slots.size());
            admit.assign(s->queue.v.begin() + (ptrdiff_t)s->queue.head,
                         s->queue.v.begin() + (ptrdiff_t)(s->queue.head + take));
            s->queue.head += take;
            s->queue.compact();
            s->st.queued = s->queue.size();
            s->queued.store(s->st.queued, std::memory_order_release);
            s->engine_resident[eng] += take;
            s->st.resident = s->engine_resident[0] + s->engine_resident[1];
            s->st.max_resident = std::max<uint64_t>(s->st.max_resident, s->st.resident);
            if (eng == 0 && !s->queue.empty()) s->cv_help.notify_one();   // the other engine has room
        }
        t_q1 = now_ns();
        for (const Query& q : admit) {
            const uint32_t i = L.free_slots.take();
            L.info[i] = SlotInfo{q.id, q.n, q.submit_ns, 0};
            L.adm[i] = q.n;
            L.admitted.push_back(i);
            L.hi = std::max(L.hi, i + 1);
        }
        L.resident += (uint32_t)admit.size();
        const uint64_t t_a = now_ns();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->st.queue_ns += t_q1 - t_q0;
            s->st.admit_ns += t_a - t_q1;
        }
        if (L.resident == 0) continue;
        const StepSignal sg{L.d_blocks, L.h_seq, L.seq + 1};
        const hipError_t e = launch_serve_step(L.stream, s->T, L.adm, L.d_slots, L.hi, s->bound, s->chunks, L.done, sg);
        if (e != hipSuccess) {
            engine_fail(s, std::string("serving engine: ") + hipGetErrorString(e));
            return;
        }
        L.seq += 1;
        L.in_flight = true;
        L.launched_ns = t_a;
    }
    // stopping: let the steps in flight finish before the pinned words go away
    for (Lane& L : E.lanes)
        if (L.in_flight) (void)hipStreamSynchronize(L.stream);
}

void free_server(gb_server* s) {
    for (Engine& E : s->engines)
        for (Lane& L : E.lanes)
            if (L.stream) (void)hipStreamSynchronize(L.stream);
    for (Engine& E : s->engines) {
        for (Lane& L : E.lanes) {
            if (L.adm) (void)hipHostFree(L.adm);
            if (L.done) (void)hipHostFree(L.done);
            if (L.h_seq) (void)hipHostFree(L.h_seq);
        }
        for (Lane& L : E.lanes)
            if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    if (s->d_slots) dfree(s->ctx, s->d_slots);
    if (s->d_blocks) dfree(s->ctx, s->d_blocks);
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
    // drained result blocks are kept, not freed, so a server that has seen a burst serves the next
    // without first-touch page faults on fresh result memory (a lane's block holds up to its slots)
    s->pool_max = std::clamp<size_t>((64ull << 20) / (((slots + 3) / 4) * sizeof(gb_result)), 4, 1024);
    s->st.slots = s->slots;
    hipError_t e = dalloc(c, &s->d_slots, slots);   // the context's budget counts the slot state
    if (e == hipSuccess) e = dalloc(c, &s->d_blocks, 4);
    // four lanes of (nearly) equal size, two per engine, laid out in the slot table in order
    uint32_t first = 0;
    release_plain2_stream(c);   // an idle plain stream would share a hardware queue with a lane
    for (uint32_t q = 0; q < 4; ++q) {
        Engine& E = s->engines[q / 2];
        Lane& L = E.lanes[q % 2];
        L.size = (uint32_t)((slots + 3 - q) / 4);
        const size_t n = std::max<uint32_t>(L.size, 1);
        if (e == hipSuccess) e = hipStreamCreate(&L.stream);
        if (e == hipSuccess) L.d_blocks = s->d_blocks + q;
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L.h_seq), sizeof(uint64_t), hipHostMallocDefault);
        if (e == hipSuccess) *L.h_seq = 0;
        if (e == hipSuccess) L.d_slots = s->d_slots + first;
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L.adm), n * sizeof(uint64_t), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&L.done), n * sizeof(ServeDone), hipHostMallocDefault);
        if (e == hipSuccess) std::memset(L.adm, 0, n * sizeof(uint64_t));
        L.info.assign(L.size, SlotInfo{0, 0, 0, 0});
        L.free_slots = FreeSlots(L.size);
        first += L.size;
    }
    if (e == hipSuccess) e = hipMemsetAsync(s->d_slots, 0, slots * sizeof(ServeSlot), s->engines[0].lanes[0].stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->d_blocks, 0, 4 * sizeof(uint32_t), s->engines[0].lanes[0].stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->engines[0].lanes[0].stream);
    if (e != hipSuccess) {
        const int rc = hip_fail(c, e, __FILE__, __LINE__);
        free_server(s);
        c->err = "serving: set-up failed -- " + c->err;
        return rc;
    }
    try {
        for (uint32_t g = 0; g < 2; ++g)
            if (s->engines[g].lanes[0].size) s->engines[g].thread = std::thread(engine_main, s, g);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->stopping.store(true);
        }
        s->cv_work.notify_all();
        s->cv_help.notify_all();
        for (Engine& E : s->engines)
            if (E.thread.joinable()) E.thread.join();
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
        s->queued.store(s->st.queued, std::memory_order_release);
        s->cv_work.notify_one();
    }
    return GB_OK;
}

int gb_serve_poll(gb_server* s, gb_result* out, uint32_t max, uint32_t* got, uint32_t wait_ms) {
    if (!s) return GB_EINVAL;
    if (got) *got = 0;
    if (max && !out) return GB_EINVAL;
    // pollers take the engines' finished blocks whole (O(1) under the engines' lock) and copy
    // them out under their own lock, so an engine never waits for a poller's copy
    std::lock_guard<std::mutex> pk(s->poll_mu);
    bool failed = false;
    uint32_t k = 0;
    for (bool first = true;; first = false) {
        if (s->polled_head == s->polled.size()) {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->polled.capacity() && s->pool.size() < s->pool_max) {
                s->polled.clear();
                s->pool.push_back(std::move(s->polled));
            }
            s->polled = std::vector<gb_result>();
            s->polled_head = 0;
            if (first && k == 0 && wait_ms && s->done.empty() && !s->failed)
                s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty() || s->failed; });
            failed = s->failed;
            if (s->done.empty()) break;
            s->polled.swap(s->done.front());
            s->done.pop_front();
        }
        const uint32_t c = (uint32_t)std::min<size_t>(max - k, s->polled.size() - s->polled_head);
        if (c) std::memcpy(out + k, s->polled.data() + s->polled_head, c * sizeof(gb_result));
        s->polled_head += c;
        k += c;
        if (k == max) break;
    }
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
    for (Engine& E : s->engines)
        if (E.thread.joinable()) E.thread.join();
    gb_ctx* c = s->ctx;
    (void)hipSetDevice(s->device);
    free_server(s);
    c->servers.fetch_sub(1);
    return GB_OK;
}

const char* gb_serve_last_error(const gb_server* s) {
    if (!s) return "";
    // an engine may set the message meanwhile: the caller gets this thread's copy, valid until
    // its next call
    thread_local std::string copy;
    std::lock_guard<std::mutex> lk(const_cast<gb_server*>(s)->mu);
    copy = s->err;
    return copy.c_str();
}

}  // extern "C"
