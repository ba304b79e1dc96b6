// gb_runlog.cpp -- resumable range verification over a run log.
//
// The reference resumes a long run by hand: the user reads where the last run
// got to and restarts with --start=N (goldbach.cu:581,604-609,689), trusting
// that every n below N was verified.  Here the range is cut into chunks of
// chunk_evens even n; each finished chunk appends one checksummed text record
// to the log (fsync'd before the next chunk starts), and a later call with the
// same log skips the chunks it already holds.  A torn or corrupt tail (a crash
// mid-write) is dropped and recomputed; a log written for different
// parameters is refused.
//
// Format (one line each, CRC-32/IEEE of the text before " crc=" in hex):
//   goldbach-run-log 1 start=S limit=L p_small=P phase2_limit=Q chunk_evens=C pmin=F crc=XXXXXXXX
//   chunk=I first=A last=B evens=E phase2=X failures=Y first_fail=Z sum_pmin=U max_pmin=V argmax=W crc=XXXXXXXX
#include <fcntl.h>
#include <sys/file.h>
#include <unistd.h>

#include <array>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gb_ctx.hpp"

using namespace gbh;

namespace {

constexpr uint64_t DEFAULT_CHUNK_EVENS = 1ull << 36;

uint32_t crc32_ieee(const char* s, size_t n) {
    // reflected polynomial 0xEDB88320, one table entry per byte value (built once, thread-safe)
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[i] = c;
        }
        return t;
    }();
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ (uint8_t)s[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

std::string with_crc(const std::string& body) {
    char tail[24];
    std::snprintf(tail, sizeof tail, " crc=%08" PRIx32 "\n", crc32_ieee(body.data(), body.size()));
    return body + tail;
}

std::string header_body(const gb_log_header& h) {
    char b[320];
    std::snprintf(b, sizeof b,
                  "goldbach-run-log 1 start=%" PRIu64 " limit=%" PRIu64 " p_small=%" PRIu64 " phase2_limit=%" PRIu64
                  " chunk_evens=%" PRIu64 " pmin=%" PRIu32,
                  h.start, h.limit, h.p_small, h.phase2_limit, h.chunk_evens, h.want_pmin);
    return b;
}

std::string record_body(const gb_log_record& r) {
    char b[400];
    std::snprintf(b, sizeof b,
                  "chunk=%" PRIu64 " first=%" PRIu64 " last=%" PRIu64 " evens=%" PRIu64 " phase2=%" PRIu64
                  " failures=%" PRIu64 " first_fail=%" PRIu64 " sum_pmin=%" PRIu64 " max_pmin=%" PRIu64
                  " argmax=%" PRIu64,
                  r.chunk, r.first_n, r.last_n, r.evens, r.phase2, r.failures, r.first_fail_n, r.sum_pmin,
                  r.max_pmin, r.argmax_n);
    return b;
}

// Splits "body crc=XXXXXXXX" and checks the checksum; false for anything else.
bool check_line(const std::string& line, std::string* body) {
    const size_t k = line.rfind(" crc=");
    if (k == std::string::npos || line.size() != k + 13) return false;
    uint32_t want = 0;
    for (size_t i = k + 5; i < line.size(); ++i) {
        const char ch = line[i];
        uint32_t d;
        if (ch >= '0' && ch <= '9') d = ch - '0';
        else if (ch >= 'a' && ch <= 'f') d = ch - 'a' + 10;
        else return false;
        want = want << 4 | d;
    }
    if (crc32_ieee(line.data(), k) != want) return false;
    *body = line.substr(0, k);
    return true;
}

bool parse_header(const std::string& body, gb_log_header* h) {
    gb_log_header t{};
    int used = 0;
    if (std::sscanf(body.c_str(),
                    "goldbach-run-log 1 start=%" SCNu64 " limit=%" SCNu64 " p_small=%" SCNu64 " phase2_limit=%" SCNu64
                    " chunk_evens=%" SCNu64 " pmin=%" SCNu32 "%n",
                    &t.start, &t.limit, &t.p_small, &t.phase2_limit, &t.chunk_evens, &t.want_pmin, &used) != 6 ||
        (size_t)used != body.size())
        return false;
    if (header_body(t) != body) return false;   // canonical form only
    *h = t;
    return true;
}

bool parse_record(const std::string& body, gb_log_record* r) {
    gb_log_record t{};
    int used = 0;
    if (std::sscanf(body.c_str(),
                    "chunk=%" SCNu64 " first=%" SCNu64 " last=%" SCNu64 " evens=%" SCNu64 " phase2=%" SCNu64
                    " failures=%" SCNu64 " first_fail=%" SCNu64 " sum_pmin=%" SCNu64 " max_pmin=%" SCNu64
                    " argmax=%" SCNu64 "%n",
                    &t.chunk, &t.first_n, &t.last_n, &t.evens, &t.phase2, &t.failures, &t.first_fail_n, &t.sum_pmin,
                    &t.max_pmin, &t.argmax_n, &used) != 10 ||
        (size_t)used != body.size())
        return false;
    if (record_body(t) != body) return false;
    *r = t;
    return true;
}

uint64_t chunk_count(const gb_log_header& h) {
    const uint64_t evens = (h.limit - h.start) / 2 + 1;
    return evens / h.chunk_evens + (evens % h.chunk_evens != 0);
}

// Bounds of chunk i: [first, last].
void chunk_bounds(const gb_log_header& h, uint64_t i, uint64_t* first, uint64_t* last) {
    *first = h.start + 2 * i * h.chunk_evens;
    const uint64_t room = (h.limit - *first) / 2;       // evens after first
    *last = *first + 2 * (room < h.chunk_evens - 1 ? room : h.chunk_evens - 1);
}

bool header_sane(const gb_log_header& h) {
    return h.start >= 4 && !(h.start & 1) && !(h.limit & 1) && h.start <= h.limit && h.chunk_evens > 0 &&
           h.want_pmin <= 1;
}

// Reads the log: header, then the longest prefix of records that are intact, in order and
// where the header's chunking puts them.  *valid_bytes: the length of that prefix in the file.
int read_log(const char* path, bool* exists, gb_log_header* h, std::vector<gb_log_record>* recs,
             uint64_t* valid_bytes, uint64_t* file_bytes, std::string* why) {
    *exists = false;
    recs->clear();
    *valid_bytes = *file_bytes = 0;
    FILE* f = std::fopen(path, "rb");
    if (!f) return GB_OK;                       // no log yet
    std::string data;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) data.append(buf, got);
    const bool bad = std::ferror(f);
    std::fclose(f);
    if (bad) { *why = "cannot read the run log"; return GB_EIO; }
    *file_bytes = data.size();
    if (data.empty()) return GB_OK;             // an empty file is a fresh log
    *exists = true;
    size_t pos = data.find('\n');
    std::string body;
    if (pos == std::string::npos || !check_line(data.substr(0, pos), &body) || !parse_header(body, h) ||
        !header_sane(*h)) {
        *why = "the run log's header is damaged or not a goldbach run log";
        return GB_EINVAL;
    }
    ++pos;
    *valid_bytes = pos;
    const uint64_t nch = chunk_count(*h);
    while (pos < data.size()) {
        const size_t e = data.find('\n', pos);
        if (e == std::string::npos) break;      // torn last line
        gb_log_record r{};
        if (!check_line(data.substr(pos, e - pos), &body) || !parse_record(body, &r)) break;
        uint64_t a, b;
        if (r.chunk != recs->size() || r.chunk >= nch) break;
        chunk_bounds(*h, r.chunk, &a, &b);
        // a chunk holding a failure may stop short of its end (the range stops at the
        // first failing segment or sub-chunk); every other chunk covers all its evens
        const uint64_t full = (b - a) / 2 + 1;
        if (r.first_n != a || r.last_n != b) break;
        if (r.failures ? (r.evens == 0 || r.evens > full) : r.evens != full) break;
        recs->push_back(r);
        pos = e + 1;
        *valid_bytes = pos;
        if (r.failures) break;                  // a failure ends the run: nothing after it counts
    }
    return GB_OK;
}

void merge(gb_seg_stats* t, const gb_log_record& r) {
    t->evens += r.evens;
    t->phase2 += r.phase2;
    if (r.failures && !t->failures) t->first_fail_n = r.first_fail_n;
    t->failures += r.failures;
    t->sum_pmin += r.sum_pmin;
    if (r.max_pmin > t->max_pmin) {             // chunks ascend: ties keep the smaller n
        t->max_pmin = r.max_pmin;
        t->argmax_n = r.argmax_n;
    }
}

int write_all(int fd, const std::string& s) {
    size_t off = 0;
    while (off < s.size()) {
        const ssize_t w = ::write(fd, s.data() + off, s.size() - off);
        if (w < 0) return -1;
        off += (size_t)w;
    }
    return 0;
}

}  // namespace

extern "C" {

int gb_log_read(const char* path, gb_log_header* header, gb_log_record* records, uint64_t cap, uint64_t* count,
                uint32_t* torn) {
    if (!path || !header || !count) return GB_EINVAL;
    bool exists = false;
    std::vector<gb_log_record> recs;
    uint64_t valid = 0, total = 0;
    std::string why;
    *header = gb_log_header{};
    const int rc = read_log(path, &exists, header, &recs, &valid, &total, &why);
    if (rc) return rc;
    if (!exists) return GB_EIO;
    *count = recs.size();
    for (uint64_t i = 0; i < recs.size() && i < cap && records; ++i) records[i] = recs[i];
    if (torn) *torn = valid != total;
    return GB_OK;
}

int gb_verify_range_logged(gb_ctx* c, uint64_t start, uint64_t limit, uint64_t p_small, const gb_tuning* tun,
                           const char* path, uint64_t chunk_evens, uint64_t max_chunks, gb_seg_stats* out,
                           gb_log_info* info, volatile uint64_t* progress) {
    if (!c) return GB_EINVAL;
    c->err.clear();
    if (!path || !*path) return fail(c, GB_EINVAL, "run log path is empty");
    int rc = normalise_range(c, &start, &limit);
    if (rc) return rc;
    const uint32_t flags = tun ? tun->flags : 0;
    gb_log_header want{};
    want.start = start;
    want.limit = limit;
    want.p_small = p_small;
    want.phase2_limit = c->cfg.phase2_limit;
    want.chunk_evens = chunk_evens ? chunk_evens : DEFAULT_CHUNK_EVENS;
    want.want_pmin = (flags & GB_WANT_PMIN) ? 1 : 0;

    // one writer at a time: the lock is held from the read to the last record
    const int fd = ::open(path, O_WRONLY | O_CREAT, 0644);
    if (fd < 0) return fail(c, GB_EIO, std::string("cannot open run log ") + path + ": " + std::strerror(errno));
    if (::flock(fd, LOCK_EX | LOCK_NB) != 0) {
        ::close(fd);
        return fail(c, GB_EAGAIN, std::string("run log ") + path + " is in use by another run");
    }
    auto io_fail = [&](const char* what) {
        const std::string m = std::string(what) + " " + path + ": " + std::strerror(errno);
        ::close(fd);
        return fail(c, GB_EIO, m);
    };

    bool exists = false;
    gb_log_header have{};
    std::vector<gb_log_record> recs;
    uint64_t valid = 0, total = 0;
    std::string why;
    rc = read_log(path, &exists, &have, &recs, &valid, &total, &why);
    if (rc) {
        ::close(fd);
        return fail(c, rc, std::string(path) + ": " + why);
    }
    if (exists && header_body(have) != header_body(want)) {
        ::close(fd);
        return fail(c, GB_EINVAL, std::string(path) + " was written for another run (" + header_body(have) +
                                      "); this run is (" + header_body(want) + ")");
    }
    // keep the intact prefix (drops a torn or corrupt tail), then append
    if (::ftruncate(fd, (off_t)valid) != 0) return io_fail("cannot truncate run log");
    if (::lseek(fd, 0, SEEK_END) < 0) return io_fail("cannot seek run log");
    if (!exists) {
        if (write_all(fd, with_crc(header_body(want))) != 0 || ::fsync(fd) != 0)
            return io_fail("cannot write run log");
    }

    gb_seg_stats tot{};
    for (const auto& r : recs) merge(&tot, r);
    const uint64_t nch = chunk_count(want);
    gb_log_info li{};
    li.resumed_chunks = recs.size();
    li.resumed_evens = tot.evens;
    li.total_chunks = nch;
    li.truncated = valid != total;
    if (progress) *progress = tot.evens;
    uint64_t ran = 0;
    double ms = 0;
    uint64_t next = recs.size();
    while (next < nch && !tot.failures && (max_chunks == 0 || ran < max_chunks)) {
        gb_log_record r{};
        r.chunk = next;
        chunk_bounds(want, next, &r.first_n, &r.last_n);
        gb_seg_stats st{};
        rc = gb_verify_range(c, r.first_n, r.last_n, p_small, tun, &st, nullptr);
        if (rc) {
            ::close(fd);
            return rc;                          // c->err holds the message; the log keeps what finished
        }
        r.evens = st.evens;
        r.phase2 = st.phase2;
        r.failures = st.failures;
        r.first_fail_n = st.first_fail_n;
        r.sum_pmin = st.sum_pmin;
        r.max_pmin = st.max_pmin;
        r.argmax_n = st.argmax_n;
        if (write_all(fd, with_crc(record_body(r))) != 0 || ::fsync(fd) != 0)
            return io_fail("cannot append to run log");
        merge(&tot, r);
        ms += st.kernel_ms;
        ++next;
        ++ran;
        if (progress) *progress = tot.evens;
    }
    ::close(fd);
    tot.kernel_ms = ms;
    li.done_chunks = next;
    li.complete = (next == nch || tot.failures) ? 1 : 0;
    if (out) *out = tot;
    if (info) *info = li;
    return GB_OK;
}

}  // extern "C"
