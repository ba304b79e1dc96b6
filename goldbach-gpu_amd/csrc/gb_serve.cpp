// gb_serve.cpp -- the continuous-batching serving engine of goldbach_gpu.h.
//
// One engine thread per server owns a HIP stream and a slot table in pinned host
// memory that the step kernel reads and writes directly (a few bytes per query,
// so no copy commands sit between a launch and its answers).
// Each iteration (a "step") it
//   1. moves queued queries into free slots (at most `slots` resident: the
//      slot count is what the memory budget allows),
//   2. launches serve_step_kernel over every resident query, new and
//      in-flight alike (one wave per query, up to chunks_per_step chunks of
//      64 candidate primes from the query's saved position),
//   3. retires the queries that finished and keeps the rest resident with
//      their new position.
// Submitters and pollers only touch host queues under a mutex, so arrivals
// overlap with the step running on the device and join at the next step.
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
constexpr uint64_t SLOT_BYTES = sizeof(ServeIn) + sizeof(ServeOut);   // slot state (pinned, device-visible)

uint64_t now_ns() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

struct Query {
    uint64_t id;
    uint64_t n;
    uint64_t k;          // next candidate index of the scan
    uint64_t submit_ns;
    uint32_t steps;
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
    hipStream_t stream = nullptr;
    ServeIn* h_in = nullptr;    // slot state in pinned host memory, read and written by the step kernel
    ServeOut* h_out = nullptr;
    std::mutex mu;
    std::condition_variable cv_work;   // engine waits for queries or stop
    std::condition_variable cv_done;   // pollers wait for results
    std::deque<Query> queue;
    std::deque<gb_result> done;
    uint64_t next_id = 0;
    std::atomic<uint64_t> arrivals{0};   // submissions so far: read without the lock by an idle engine
    std::atomic<bool> stopping{false};
    bool stop = false;
    bool failed = false;
    std::string err;
    gb_serve_stats st{};
    std::thread engine;
};

namespace {

void engine_fail(gb_server* s, const std::string& msg) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->failed = true;
    s->err = msg;
    s->cv_done.notify_all();
}

void engine_main(gb_server* s) {
    if (hipSetDevice(s->device) != hipSuccess) {
        engine_fail(s, "serving engine: hipSetDevice failed");
        return;
    }
    std::vector<Query> active;
    active.reserve(s->slots);
    uint64_t seen = 0, left = 0;   // submissions counted and queries left queued at the last admission
    for (;;) {
        if (active.empty() && left == 0) {
            // idle: poll for arrivals for a short while, so a query that arrives now does not
            // wait for the thread to be woken; then sleep on the condition variable
            const uint64_t t0 = now_ns();
            while (s->arrivals.load(std::memory_order_acquire) == seen && !s->stopping.load(std::memory_order_acquire) &&
                   now_ns() - t0 < IDLE_SPIN_NS) {
            }
        }
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !active.empty(); });
            seen = s->arrivals.load(std::memory_order_relaxed);
            if (s->stop) return;
            while (active.size() < s->slots && !s->queue.empty()) {
                active.push_back(s->queue.front());
                s->queue.pop_front();
            }
            s->st.queued = s->queue.size();
            s->st.resident = active.size();
            left = s->queue.size();
        }
        const uint32_t m = (uint32_t)active.size();
        for (uint32_t i = 0; i < m; ++i) s->h_in[i] = ServeIn{active[i].n, active[i].k};
        // the step reads its inputs from and writes its results to the pinned staging directly:
        // a few bytes per query, so no copy commands sit between the launch and the answer
        hipError_t e = launch_serve_step(s->stream, s->T, s->h_in, m, s->bound, s->chunks, s->h_out);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            engine_fail(s, std::string("serving engine: ") + hipGetErrorString(e));
            return;
        }
        const uint64_t t = now_ns();
        uint32_t keep = 0;
        std::lock_guard<std::mutex> lk(s->mu);
        for (uint32_t i = 0; i < m; ++i) {
            Query q = active[i];
            q.steps += 1;
            if (s->h_out[i].state) {
                gb_result r;
                r.id = q.id;
                r.n = q.n;
                r.p = s->h_out[i].p;
                r.submit_ns = q.submit_ns;
                r.done_ns = t;
                r.steps = q.steps;
                r.reserved = 0;
                s->done.push_back(r);
            } else {
                q.k = s->h_out[i].k;
                active[keep++] = q;
            }
        }
        active.resize(keep);
        s->st.steps += 1;
        s->st.step_queries += m;
        s->st.max_resident = std::max<uint64_t>(s->st.max_resident, m);
        s->st.completed += m - keep;
        s->st.resident = keep;
        if (keep < m) s->cv_done.notify_all();
    }
}

void free_server(gb_server* s) {
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->h_in) (void)hipHostFree(s->h_in);
    if (s->h_out) (void)hipHostFree(s->h_out);
    if (s->stream) (void)hipStreamDestroy(s->stream);
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
    if (k.mem_budget) slots = std::min<uint64_t>(slots, k.mem_budget / SLOT_BYTES);
    if (slots == 0) return fail(c, GB_ENOMEM, "serving: the memory budget holds no slot");
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
    hipError_t e = hipStreamCreate(&s->stream);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&s->h_in), slots * sizeof(ServeIn), hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&s->h_out), slots * sizeof(ServeOut), hipHostMallocDefault);
    if (e != hipSuccess) {
        free_server(s);
        return fail(c, e == hipErrorOutOfMemory ? GB_ENOMEM : GB_EDEVICE,
                    std::string("serving: set-up failed -- ") + hipGetErrorString(e));
    }
    try {
        s->engine = std::thread(engine_main, s);
    } catch (...) {
        free_server(s);
        return fail(c, GB_EDEVICE, "serving: could not start the engine thread");
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
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed) return GB_EDEVICE;
    if (s->queue.size() + count > s->queue_cap) {
        s->err = "gb_serve_submit: queue full";
        return GB_EAGAIN;
    }
    if (first_id) *first_id = s->next_id;
    for (uint32_t i = 0; i < count; ++i) s->queue.push_back(Query{s->next_id++, n[i], 0, t, 0});
    s->arrivals.fetch_add(count, std::memory_order_release);
    s->st.submitted += count;
    s->st.queued = s->queue.size();
    if (count) s->cv_work.notify_one();
    return GB_OK;
}

int gb_serve_poll(gb_server* s, gb_result* out, uint32_t max, uint32_t* got, uint32_t wait_ms) {
    if (!s) return GB_EINVAL;
    if (got) *got = 0;
    if (max && !out) return GB_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    if (wait_ms && s->done.empty() && !s->failed)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty() || s->failed; });
    uint32_t k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    if (got) *got = k;
    if (s->failed && k == 0) return GB_EDEVICE;
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
    if (s->engine.joinable()) s->engine.join();
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
