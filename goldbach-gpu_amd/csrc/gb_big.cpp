// gb_big.cpp -- the arbitrary-size single-n search behind gb_big_min_partition
// (the reference's big_check, big_check.cpp:35-144), plus decimal conversion.
//
// The reference walks the primes p <= 10^7 in batches of 1,000 over OpenMP
// threads and reports the p of whichever thread first sees GMP accept
// n - p (so p is near-minimal).  Here:
//   1. the device computes n mod ell for the odd primes ell < 2^20 and marks
//      every odd p <= p_max with ell | n - p (such q are composite: q > ell);
//   2. the surviving primes p, ascending, are tested in batches of one
//      workgroup each (base 2);
//   3. the smallest p of a batch whose q passed is confirmed with the other
//      bases in one more launch (one workgroup per base).
// Every smaller p failed a test or has a small factor in n - p, so the p
// returned is the smallest whose n - p passes all `rounds` bases.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gb_big.hpp"
#include "gb_ctx.hpp"

using namespace gbh;

namespace {

constexpr uint32_t FILTER_BOUND = 1u << 20;          // small prime factors excluded before testing
constexpr uint64_t DEFAULT_P_MAX = 10'000'000ull;    // big_check.cpp:58
constexpr uint32_t MAX_ROUNDS = 25;
constexpr uint32_t BASES[MAX_ROUNDS] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29, 31, 37, 41,
                                        43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97};

// Odd primes in [3, limit] by an odd-only byte sieve.
std::vector<uint32_t> odd_primes(uint64_t limit) {
    std::vector<uint32_t> out;
    if (limit < 3) return out;
    const uint64_t nb = (limit - 1) / 2;             // index i <-> 2i + 1, i in [1, nb]
    std::vector<uint8_t> comp(nb + 1, 0);
    for (uint64_t i = 1; (2 * i + 1) * (2 * i + 1) <= limit; ++i) {
        if (comp[i]) continue;
        const uint64_t p = 2 * i + 1;
        for (uint64_t j = (p * p - 1) / 2; j <= nb; j += p) comp[j] = 1;
    }
    for (uint64_t i = 1; i <= nb; ++i)
        if (!comp[i]) out.push_back((uint32_t)(2 * i + 1));
    return out;
}

uint32_t trimmed(const uint32_t* w, uint32_t n) {
    while (n > 1 && w[n - 1] == 0) --n;
    return n;
}

// A device buffer for one call, counted against the context's memory budget.
template <typename T>
struct DevBuf {
    gb_ctx* c;
    T* p = nullptr;
    explicit DevBuf(gb_ctx* ctx) : c(ctx) {}
    ~DevBuf() { dfree(c, p); }
    hipError_t alloc(size_t count) { return dalloc(c, &p, count); }
};

}  // namespace

extern "C" {

int gb_big_from_decimal(const char* s, uint32_t* words, uint32_t cap, uint32_t* nwords) {
    if (!s || !words || !nwords || cap == 0) return GB_EINVAL;
    const size_t len = std::strlen(s);
    if (len == 0) return GB_EINVAL;
    for (size_t i = 0; i < len; ++i)
        if (s[i] < '0' || s[i] > '9') return GB_EINVAL;
    uint32_t n = 1;
    words[0] = 0;
    // n = n * 10^k + chunk over chunks of up to 9 digits
    for (size_t i = 0; i < len;) {
        const size_t k = std::min<size_t>(9, len - i);
        uint32_t mul = 1, chunk = 0;
        for (size_t j = 0; j < k; ++j) {
            mul *= 10;
            chunk = chunk * 10 + (uint32_t)(s[i + j] - '0');
        }
        uint64_t c = chunk;
        for (uint32_t j = 0; j < n; ++j) {
            const uint64_t v = (uint64_t)words[j] * mul + c;
            words[j] = (uint32_t)v;
            c = v >> 32;
        }
        if (c) {
            if (n == cap) return GB_ERANGE;
            words[n++] = (uint32_t)c;
        }
        i += k;
    }
    *nwords = trimmed(words, n);
    return GB_OK;
}

int gb_big_to_decimal(const uint32_t* words, uint32_t nwords, char* out, size_t cap) {
    if (!words || !out || nwords == 0) return GB_EINVAL;
    std::vector<uint32_t> w(words, words + trimmed(words, nwords));
    std::vector<uint32_t> chunks;                     // base 10^9, least significant first
    while (!(w.size() == 1 && w[0] == 0)) {
        uint64_t r = 0;
        for (size_t j = w.size(); j-- > 0;) {
            const uint64_t v = (r << 32) | w[j];
            w[j] = (uint32_t)(v / 1000000000u);
            r = v % 1000000000u;
        }
        chunks.push_back((uint32_t)r);
        while (w.size() > 1 && w.back() == 0) w.pop_back();
    }
    std::string s = chunks.empty() ? "0" : std::to_string(chunks.back());
    for (size_t i = chunks.size(); i-- > 1;) {
        char b[16];
        std::snprintf(b, sizeof b, "%09u", chunks[i - 1]);
        s += b;
    }
    if (s.size() + 1 > cap) return GB_ERANGE;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return GB_OK;
}

int gb_big_min_partition(gb_ctx* c, const uint32_t* n_words, uint32_t nwords, uint32_t rounds, uint64_t p_max,
                         uint64_t* p_out, gb_big_stats* st) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!n_words || nwords == 0 || !p_out) return fail(c, GB_EINVAL, "gb_big_min_partition: null argument");
    *p_out = 0;
    gb_big_stats stats{};
    const uint32_t L = trimmed(n_words, nwords);
    stats.words = L;
    if (n_words[0] & 1) return fail(c, GB_EINVAL, "n must be even");
    if (L == 1 && n_words[0] < 4) return fail(c, GB_EINVAL, "n must be >= 4");
    if (rounds == 0) rounds = MAX_ROUNDS;
    if (rounds > MAX_ROUNDS) return fail(c, GB_EINVAL, "rounds must be <= 25");
    if (p_max == 0) p_max = DEFAULT_P_MAX;
    if (p_max > 0xFFFFFFFFull) return fail(c, GB_ERANGE, "p_max must be below 2^32");
    if (L <= 2) {
        // below 2^64 the strong test to bases 2..37 is exact: the 64-bit path gives the same p
        const uint64_t n = (uint64_t)n_words[0] | (L == 2 ? (uint64_t)n_words[1] << 32 : 0);
        uint64_t p = 0;
        const int rc = min_partition_bounded(c, &n, 1, p_max, &p);
        if (rc) return rc;
        *p_out = p;
        if (st) *st = stats;
        return GB_OK;
    }
    if (L > gbk::BIG_LMAX) return fail(c, GB_ERANGE, "n has more than 2304 32-bit words (73,728 bits)");
    HCK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;

    const std::vector<uint32_t> primes = odd_primes(p_max);
    const std::vector<uint32_t> ell = odd_primes(FILTER_BOUND);
    const uint64_t bad_words = p_max / 64 + 1;
    DevBuf<uint32_t> d_n(c), d_ell(c), d_r(c), d_bad(c), d_pass(c);
    DevBuf<gbk::BigTask> d_tasks(c);
    // one batch = the tests the device holds at once (LDS and 2048 threads per CU); a
    // 1024-thread test already fills its CU (two per CU took twice as long as one), so those
    // run one per CU and the batches stay short
    const size_t lds = gbk::big_lds_bytes(L) + 256;
    const uint32_t threads = gbk::big_threads(L);
    const uint32_t per_cu = threads >= 1024 ? 1u
                                            : std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)(gbk::LDS_PER_CU / lds),
                                                                                       2048 / threads));
    const uint32_t batch = (uint32_t)std::max(1, c->cus) * per_cu;
    if (d_n.alloc(L) != hipSuccess || d_ell.alloc(ell.size()) != hipSuccess || d_r.alloc(ell.size()) != hipSuccess ||
        d_bad.alloc(bad_words) != hipSuccess || d_tasks.alloc(std::max<uint32_t>(batch, MAX_ROUNDS)) != hipSuccess ||
        d_pass.alloc(std::max<uint32_t>(batch, MAX_ROUNDS)) != hipSuccess)
        return hip_fail(c, hipErrorOutOfMemory, __FILE__, __LINE__);
    HCK(c, hipMemcpyAsync(d_n.p, n_words, L * 4, hipMemcpyHostToDevice, s));
    HCK(c, hipMemcpyAsync(d_ell.p, ell.data(), ell.size() * 4, hipMemcpyHostToDevice, s));
    HCK(c, hipMemsetAsync(d_bad.p, 0, bad_words * 4, s));
    HCK(c, hipEventRecord(c->ev0, s));
    const uint32_t nsmall = (uint32_t)(std::lower_bound(ell.begin(), ell.end(), 1u << 13) - ell.begin());
    HCK(c, gbk::launch_big_filter(s, d_n.p, L, d_ell.p, (uint32_t)ell.size(), nsmall, d_r.p, p_max, d_bad.p));
    HCK(c, hipEventRecord(c->ev1, s));
    std::vector<uint32_t> bad(bad_words);
    HCK(c, hipMemcpyAsync(bad.data(), d_bad.p, bad_words * 4, hipMemcpyDeviceToHost, s));
    HCK(c, hipStreamSynchronize(s));
    float ms = 0;
    HCK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    stats.kernel_ms += ms;

    std::vector<uint32_t> cand;                       // primes p whose n - p has no factor < 2^20
    std::vector<uint32_t> rank;                       // index of each candidate among the primes
    cand.reserve(primes.size() / 8);
    for (size_t i = 0; i < primes.size(); ++i) {
        const uint32_t p = primes[i];
        if (!((bad[p >> 6] >> ((p >> 1) & 31)) & 1)) {
            cand.push_back(p);
            rank.push_back((uint32_t)i);
        }
    }
    std::vector<gbk::BigTask> tasks(std::max<uint32_t>(batch, MAX_ROUNDS));
    std::vector<uint32_t> pass(tasks.size());
    auto run = [&](uint32_t count) -> int {
        HCK(c, hipMemcpyAsync(d_tasks.p, tasks.data(), count * sizeof(gbk::BigTask), hipMemcpyHostToDevice, s));
        HCK(c, hipEventRecord(c->ev0, s));
        HCK(c, gbk::launch_big_mr(s, d_n.p, L, d_tasks.p, count, d_pass.p));
        HCK(c, hipEventRecord(c->ev1, s));
        HCK(c, hipMemcpyAsync(pass.data(), d_pass.p, count * 4, hipMemcpyDeviceToHost, s));
        HCK(c, hipStreamSynchronize(s));
        float t = 0;
        HCK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
        stats.kernel_ms += t;
        stats.mr_tasks += count;
        ++stats.batches;
        return GB_OK;
    };
    // A candidate (n - p free of factors below 2^20) is prime with probability about
    // 1 / (ln n * prod_(ell < 2^20) (1 - 1/ell)) = 1 / (0.0405 ln n): the first batch takes three
    // times that many (a 5 % chance of needing a second), at least one per CU, at most a full batch.
    const double ln_n = (32.0 * (L - 1) + std::log2((double)n_words[L - 1] + 1.0)) * 0.6931471805599453;
    const uint32_t first =
        (uint32_t)std::min<double>(batch, std::max<double>(std::max(1, c->cus), 3.0 * 0.0405 * ln_n));
    uint64_t found = 0;
    size_t next = 0;
    while (!found && next < cand.size()) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(next == 0 ? first : batch, cand.size() - next);
        for (uint32_t i = 0; i < cnt; ++i) tasks[i] = gbk::BigTask{cand[next + i], 2};
        int rc = run(cnt);
        if (rc) return rc;
        const std::vector<uint32_t> first(pass.begin(), pass.begin() + cnt);
        for (uint32_t i = 0; i < cnt && !found; ++i) {
            if (!first[i]) continue;
            bool all = true;
            if (rounds > 1) {
                for (uint32_t k = 1; k < rounds; ++k) tasks[k - 1] = gbk::BigTask{cand[next + i], BASES[k]};
                rc = run(rounds - 1);
                if (rc) return rc;
                for (uint32_t k = 0; k + 1 < rounds; ++k) all = all && pass[k];
            }
            if (all) {
                found = cand[next + i];
                stats.candidates = next + i + 1;
                stats.primes_scanned = rank[next + i] + 1;
            }
        }
        next += cnt;
    }
    if (!found) {
        stats.candidates = cand.size();
        stats.primes_scanned = primes.size();
    }
    *p_out = found;
    if (st) *st = stats;
    return GB_OK;
}

}  // extern "C"
