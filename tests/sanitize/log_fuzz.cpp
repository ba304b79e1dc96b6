// Reads every log file named on the command line with gb_log_read (into a 64-record buffer,
// then counting only), for a host build of the run-log code under AddressSanitizer and
// UBSan.  The two library functions the reader's translation unit also references
// (range normalisation and the device run of a chunk) are stand-ins here: reading a log
// never calls them.
#include <cstdint>
#include <cstdio>
#include <string>

#include "goldbach_gpu.h"

struct gb_ctx;
namespace gbh {
int normalise_range(gb_ctx*, uint64_t*, uint64_t*) { return GB_EINVAL; }
int fail(gb_ctx*, int rc, const std::string&) { return rc; }
}  // namespace gbh
extern "C" int gb_verify_range(gb_ctx*, uint64_t, uint64_t, uint64_t, const gb_tuning*, gb_seg_stats*,
                               volatile uint64_t*) {
    return GB_EINVAL;
}

int main(int argc, char** argv) {
    static gb_log_record recs[64];
    long ok = 0, refused = 0;
    for (int i = 1; i < argc; ++i) {
        gb_log_header h;
        uint64_t n = 0, n2 = 0;
        uint32_t torn = 0, torn2 = 0;
        const int rc = gb_log_read(argv[i], &h, recs, 64, &n, &torn);
        const int rc2 = gb_log_read(argv[i], &h, nullptr, 0, &n2, &torn2);
        if (rc != rc2 || (rc == GB_OK && (n != n2 || torn != torn2))) {
            std::printf("inconsistent result for %s\n", argv[i]);
            return 2;
        }
        rc == GB_OK ? ++ok : ++refused;
    }
    std::printf("read %d logs: %ld accepted, %ld refused\n", argc - 1, ok, refused);
    return 0;
}
