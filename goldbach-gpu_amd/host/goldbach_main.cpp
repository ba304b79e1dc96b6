// goldbach_main.cpp -- the `goldbach` command line, drop-in for the reference's
// range verifier (src/goldbach.cu:543-801): same arguments, same checks, same
// messages and exit codes, running on one MI355X through include/goldbach_gpu.h.
//
//   goldbach <LIMIT> [SEG_SIZE] [P_SMALL]
//   goldbach <LIMIT> [--seg-size=N] [--p-small=N] [--batch-size=N] [--gpus=N]
//                    [--start=N] [--progress] [--async]
// Additions: --form=batched|plain (default batched: one persistent fused
// launch; plain: one launch sequence and host sync per segment, as the
// reference), --tile-log2=N (batched tile of 2^N evens), --phase2-limit=N
// (Phase-2 prime bound, 1e8 as in the reference; a test hook), --autotune
// (time tile / segment-size candidates on samples of the range first),
// --counts (print c(n), the number of partitions p + q = n with p <= q, for
// every even n of the range, as the reference's cpu_goldbach.cpp:101-103 does
// with PRINT_COUNTS, then the smallest c(n) seen; n at most 2^36),
// --log=FILE [--log-chunk=N] (resumable run: every N evens finished append a
// checksummed record to FILE; running the same command again continues after
// the last record instead of restarting by hand with --start).
// --batch-size and --async are accepted and have no effect: the device keeps
// every candidate prime resident, so there are no prime batches to size.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "goldbach_gpu.h"

namespace {

void print_usage(const char* prog) {
    std::cout << "Goldbach Verifier (MI355X)\n\n"
              << "Usage:\n"
              << "  " << prog << " <LIMIT> [SEG_SIZE] [P_SMALL]\n"
              << "  " << prog << " <LIMIT> [--seg-size=N] [--p-small=N] [--gpus=N]\n\n"
              << "Required:\n"
              << "  LIMIT            Max even integer to check\n\n"
              << "Optional:\n"
              << "  --seg-size=N     Even integers per segment (default: 10,000,000)\n"
              << "  --p-small=N      GPU prime search bound (max: 4,000,000,000)\n"
              << "  --batch-size=N   Accepted for compatibility (primes stay resident on the device)\n"
              << "  --gpus=N         Number of GPUs to use (default: 1)\n"
              << "  --start=N        Starting number for verification (default: 4)\n"
              << "  --progress       Show real-time progress updates\n"
              << "  --form=F         batched (default) or plain per-segment launches\n"
              << "  --tile-log2=N    Batched tile of 2^N evens (12..20; default 19 below 2^44, else 20)\n"
              << "  --phase2-limit=N Phase-2 prime bound (default: 100000000)\n"
              << "  --autotune       Pick tile / segment size by timing samples of the range\n"
              << "  --counts         Print c(n), the number of Goldbach partitions, for every n\n"
              << "  --log=FILE       Resumable run: record finished chunks in FILE, continue after them\n"
              << "  --log-chunk=N    Even integers per run-log record (default: 2^36)\n"
              << "  -h, --help       Show this help message\n";
}

double secs_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        print_usage(argv[0]);
        return 0;
    }
    uint64_t LIMIT = 0, SEG_SIZE = 10'000'000ULL, P_SMALL = 1'000'000ULL, START = 4;
    uint64_t phase2_limit = 100'000'000ULL;
    uint32_t tile_log2 = 0;
    int requested_gpus = 1;
    bool progress = false;
    bool autotune = false;
    bool counts = false;
    std::string log_path;
    uint64_t log_chunk = 0;
    uint32_t form = GB_FORM_BATCHED;
    std::vector<std::string> positional;

    try {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            auto has = [&](const char* p) { return a.rfind(p, 0) == 0; };
            if (a == "-h" || a == "--help") {
                print_usage(argv[0]);
                return 0;
            }
            if (a == "--async") continue;
            if (a == "--progress") { progress = true; continue; }
            if (a == "--autotune") { autotune = true; continue; }
            if (a == "--counts") { counts = true; continue; }
            if (has("--log=")) { log_path = a.substr(6); continue; }
            if (has("--log-chunk=")) { log_chunk = std::stoull(a.substr(12)); continue; }
            if (has("--batch-size=")) { (void)std::stoull(a.substr(13)); continue; }
            if (has("--gpus=")) { requested_gpus = std::stoi(a.substr(7)); continue; }
            if (has("--seg-size=")) { SEG_SIZE = std::stoull(a.substr(11)); continue; }
            if (has("--p-small=")) { P_SMALL = std::stoull(a.substr(10)); continue; }
            if (has("--start=")) { START = std::stoull(a.substr(8)); continue; }
            if (has("--phase2-limit=")) { phase2_limit = std::stoull(a.substr(15)); continue; }
            if (has("--tile-log2=")) { tile_log2 = (uint32_t)std::stoul(a.substr(12)); continue; }
            if (has("--form=")) {
                const std::string f = a.substr(7);
                if (f == "plain") form = GB_FORM_PLAIN;
                else if (f == "batched") form = GB_FORM_BATCHED;
                else throw std::invalid_argument("form");
                continue;
            }
            positional.push_back(a);   // an unknown --flag lands here, as in the reference
        }
        if (positional.size() >= 1) LIMIT = std::stoull(positional[0]);
        if (positional.size() >= 2) SEG_SIZE = std::stoull(positional[1]);
        if (positional.size() >= 3) P_SMALL = std::stoull(positional[2]);
    } catch (...) {
        std::cerr << "Error: Invalid numeric argument.\n";
        return 1;
    }

    if (LIMIT < 4) { std::cerr << "Error: LIMIT must be >= 4.\n"; return 1; }
    if (LIMIT % 2 != 0) LIMIT--;
    if (SEG_SIZE == 0 || SEG_SIZE % 2 != 0) { std::cerr << "Error: SEG_SIZE must be even and > 0.\n"; return 1; }
    const uint64_t MAX_P_SMALL = 4'000'000'000ULL;
    if (P_SMALL > MAX_P_SMALL) {
        std::cerr << "Error: P_SMALL must be <= " << MAX_P_SMALL << " to prevent mathematical overflow.\n";
        return 1;
    }
    if (P_SMALL > LIMIT) P_SMALL = LIMIT;
    if (START < 4) START = 4;
    if (START % 2 != 0) START++;
    if (START >= LIMIT) { std::cerr << "Error: START must be less than LIMIT.\n"; return 1; }
    if (P_SMALL < 1'000'000ULL) {
        std::cerr << "\n[!] WARNING: P_SMALL = " << P_SMALL << " is very small.\n"
                  << "    This may cause excessive Phase 2 fallbacks.\n"
                  << "    Recommended: P_SMALL >= 10^7 for numbers > 10^12\n"
                  << "                 P_SMALL >= 10^8 for numbers > 10^18\n\n";
    }

    int device_count = 0;
    if (gb_device_count(&device_count) != GB_OK || device_count == 0) {
        std::cerr << "No GPU devices found.\n";
        return 1;
    }
    if (requested_gpus > device_count)
        std::cerr << "Requested " << requested_gpus << " GPUs, but only " << device_count << " available. Using "
                  << device_count << ".\n";
    const int use_gpus = (requested_gpus <= 0 || requested_gpus > device_count) ? device_count : requested_gpus;
    if (use_gpus > 1)
        std::cout << "[Info] this build verifies on one device; using GPU 0 of " << use_gpus << ".\n";

    // exact isqrt (goldbach.cu:636-646) and the small-prime bound the reference builds
    uint64_t sqrt_limit = 0;
    {
        uint64_t lo = 1, hi = LIMIT > (1ULL << 32) ? (1ULL << 32) : LIMIT;
        while (lo <= hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (mid <= LIMIT / mid) { sqrt_limit = mid; lo = mid + 1; }
            else hi = mid - 1;
        }
    }
    uint64_t small_high = std::max(sqrt_limit + 1, P_SMALL);
    if (small_high % 2 == 0) small_high++;

    char name[256] = {0};
    uint64_t total_mem = 0;
    int cus = 0;
    gb_device_info(0, name, sizeof name, &total_mem, &cus);
    if ((SEG_SIZE + 262143) / 262144 > 0xFFFFFFFFull) {
        std::cerr << "[!] ERROR: Segment size too large. Grid dimensions exceed uint32_t limit.\n"
                  << "    Reduce SEG_SIZE or P_SMALL.\n";
        return 1;
    }

    std::cout << "Building small primes bitset up to " << small_high << " (device-resident)...\n";
    const auto t0 = std::chrono::steady_clock::now();
    gb_config cfg{};
    cfg.max_seg_evens = SEG_SIZE;
    cfg.phase2_limit = phase2_limit;
    gb_ctx* ctx = nullptr;
    const int rc_open = gb_open(0, &cfg, &ctx);
    if (rc_open == GB_ENOMEM) {
        std::cerr << "\n[!] ERROR: GPU 0 (" << name << ") has insufficient VRAM.\n"
                  << "    " << gb_last_error(nullptr) << "\n"
                  << "    Available: " << total_mem / (1024 * 1024) << " MB\n"
                  << "    Reduce SEG_SIZE or use a smaller LIMIT.\n";
        return 1;
    }
    if (rc_open != GB_OK) {
        std::cerr << "\n[!] Program aborted due to internal hardware/HIP errors.\n" << gb_last_error(nullptr) << "\n";
        return 1;
    }
    std::cout << "[Hardware] GPU 0: " << name << " (" << total_mem / (1024 * 1024) << " MB VRAM, " << cus
              << " CUs)\n";
    std::cout << "Phase-2 primes up to " << phase2_limit << " resident on the device.\n";
    std::cout << "Initialization completed in " << secs_since(t0) * 1000.0 << " ms.\n\n";

    const uint64_t total_even = (LIMIT - START) / 2 + 1;
    std::cout << "--- Launching Verifier ---\n";
    std::cout << "Checking range : [" << START << ", " << LIMIT << "]\n";
    std::cout << "Total numbers  : " << total_even << "\n\n";

    volatile uint64_t processed = 0;
    std::atomic<bool> running{true};
    std::thread progress_thread;
    const auto t_main = std::chrono::steady_clock::now();
    auto print_progress = [&](uint64_t done) {
        const double el = secs_since(t_main);
        const double pct = 100.0 * (double)done / (double)total_even;
        const double rate = el > 0 ? (double)done / el : 0.0;
        const uint64_t eta = rate > 0 ? (uint64_t)((double)(total_even - done) / rate) : 0;
        std::cout << "\r[Progress] " << done << " / " << total_even << " (" << std::fixed << std::setprecision(2)
                  << pct << "%) | " << std::scientific << std::setprecision(2) << rate << " numbers/sec | ETA: " << eta
                  << "s          " << std::defaultfloat << std::flush;
    };
    if (progress) {
        print_progress(0);
        progress_thread = std::thread([&] {
            while (running.load()) {
                std::this_thread::sleep_for(std::chrono::milliseconds(500));
                if (running.load()) print_progress(processed);
            }
        });
    }

    if (counts) {
        // c(n) in slices of 2^22 n: 32 MiB of host buffer whatever the range
        std::vector<uint64_t> c(std::min<uint64_t>(total_even, 1ull << 22));
        uint64_t c_min = ~0ull, n_min = 0;
        auto stop_progress = [&] {
            if (!progress) return;
            running.store(false);
            progress_thread.join();
            std::cout << "\n";
        };
        for (uint64_t done = 0; done < total_even;) {
            const uint64_t k = std::min<uint64_t>(c.size(), total_even - done);
            const uint64_t n0 = START + 2 * done;
            if (gb_count_partitions(ctx, n0, k, c.data()) != GB_OK) {
                stop_progress();
                std::cerr << "[!] FATAL ERROR in GPU 0 Worker: " << gb_last_error(ctx) << "\n";
                gb_close(ctx);
                return 1;
            }
            for (uint64_t i = 0; i < k; ++i) {
                std::cout << "c(" << n0 + 2 * i << ") = " << c[i] << "\n";
                if (c[i] < c_min) { c_min = c[i]; n_min = n0 + 2 * i; }
            }
            done += k;
            processed = done;
        }
        stop_progress();
        gb_close(ctx);
        std::cout << "\n--- Partition Counts Complete ---\n";
        std::cout << "Even numbers counted   : " << total_even << "\n";
        std::cout << "Smallest c(n)          : " << c_min << " at n = " << n_min << "\n";
        std::cout << "Total computation time : " << secs_since(t_main) << " seconds\n";
        return c_min == 0 ? 1 : 0;
    }

    gb_tuning tun{};
    tun.form = form;
    tun.seg_evens = SEG_SIZE;
    tun.tile_log2 = tile_log2;
    if (autotune) {
        double rate = 0;
        if (gb_autotune(ctx, START, LIMIT, P_SMALL, form, 0, &tun, &rate) != GB_OK) {
            std::cerr << "[!] FATAL ERROR in GPU 0 Worker: " << gb_last_error(ctx) << "\n";
            gb_close(ctx);
            return 1;
        }
        if (form == GB_FORM_BATCHED)
            std::cout << "Autotuned      : tile size index " << tun.tile_log2 << ", " << tun.wg_per_cu
                      << " workgroup(s) per CU, lane items below stream/"
                      << (tun.lane_item_div ? tun.lane_item_div : 128u) << ", wave items below stream/"
                      << (tun.wave_item_div ? tun.wave_item_div : 1024u);
        else
            std::cout << "Autotuned      : segments of " << tun.seg_evens << " evens";
        std::cout << " (" << std::scientific << std::setprecision(3) << rate << std::defaultfloat
                  << " evens/s on the samples)\n\n";
        // the pick is cached on the context: the range call applies it to the part of the range it
        // was timed on (a run log's chunks take it as given)
        if (log_path.empty() && form == GB_FORM_BATCHED) {
            tun.flags |= GB_TUNE_AUTO;
            tun.tile_log2 = 0;
            tun.wg_per_cu = 0;
        }
    }
    gb_seg_stats st{};
    gb_log_info li{};
    const int rc = log_path.empty()
                       ? gb_verify_range(ctx, START, LIMIT, P_SMALL, &tun, &st, &processed)
                       : gb_verify_range_logged(ctx, START, LIMIT, P_SMALL, &tun, log_path.c_str(), log_chunk, 0, &st,
                                                &li, &processed);
    const double total_ms = secs_since(t_main) * 1000.0;
    if (progress) {
        running.store(false);
        progress_thread.join();
        print_progress(rc == GB_OK ? st.evens : processed);
        std::cout << "\n";
    }
    if (rc != GB_OK) {
        std::cerr << "[!] FATAL ERROR in GPU 0 Worker: " << gb_last_error(ctx) << "\n";
        std::cerr << "\n[!] Program aborted due to internal hardware/HIP errors.\n";
        gb_close(ctx);
        return 1;
    }
    gb_close(ctx);
    if (!log_path.empty())
        std::cout << "Run log        : " << log_path << " (" << li.resumed_chunks << " of " << li.total_chunks
                  << " chunks resumed, " << li.resumed_evens << " evens" << (li.truncated ? "; damaged tail dropped" : "")
                  << ")\n";
    if (st.failures) {
        std::cout << "\n[!] Goldbach FAILED at n = " << st.first_fail_n << "\n";
        return 1;
    }
    std::cout << "\n--- Verification Complete ---\n";
    std::cout << "All even numbers from " << START << " up to " << LIMIT << " satisfy Goldbach. ✓\n";
    std::cout << "Total computation time : " << (total_ms / 1000.0) << " seconds\n";
    std::cout << "Phase 2 fallbacks      : " << st.phase2 << "\n";
    std::cout << "Device kernel time     : " << st.kernel_ms << " ms\n";
    return 0;
}
