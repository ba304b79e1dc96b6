// single_check_main.cpp -- the `single_check` command line, drop-in for the
// reference's single-number checker (src/single_check.cu:284-319): same
// argument checks and messages.  The partition printed is the minimal one
// (smallest prime p with n - p prime), found on the device by
// gb_min_partition_batch; the reference prints whichever GPU thread finished
// first (RESULTS.md:104-110 notes it is not minimal).
#include <chrono>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>

#include "goldbach_gpu.h"

int main(int argc, char** argv) {
    if (argc < 2 || std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help") {
        std::cout << "GPU Single Number Goldbach Checker (MI355X)\n";
        std::cout << "Usage: " << argv[0] << " <N>\n";
        std::cout << "Example: " << argv[0] << " 1000000000000\n";
        return 0;
    }
    uint64_t n = 0;
    try {
        n = std::stoull(argv[1]);
    } catch (const std::out_of_range&) {
        std::cerr << "Error: Number is too large for 64-bit unsigned integer.\n";
        std::cerr << "Use 'big_check' for numbers > 1.84 x 10^19.\n";
        return 1;
    } catch (const std::exception&) {
        std::cerr << "Error: Invalid numeric argument.\n";
        return 1;
    }
    if (n < 4) {
        std::cerr << "Error: n must be >= 4.\n";
        return 1;
    }
    if (n % 2 != 0) {
        std::cerr << "Error: n must be an even integer.\n";
        return 1;
    }
    std::cout << "Checking Goldbach for n = " << n << "...\n";
    const auto t0 = std::chrono::steady_clock::now();
    gb_ctx* ctx = nullptr;
    if (gb_open(0, nullptr, &ctx) != GB_OK) {
        std::cerr << "Error: " << gb_last_error(nullptr) << "\n";
        return 1;
    }
    const auto t1 = std::chrono::steady_clock::now();
    uint64_t p = 0;
    const int rc = gb_min_partition_batch(ctx, &n, 1, &p);
    const auto t2 = std::chrono::steady_clock::now();
    if (rc != GB_OK) {
        std::cerr << "Error: " << gb_last_error(ctx) << "\n";
        gb_close(ctx);
        return 1;
    }
    gb_close(ctx);
    std::cout << "\n--- Result ---\n";
    if (p) {
        std::cout << n << " = " << p << " + " << (n - p) << "\n";
        std::cout << "Goldbach holds. ✓\n";
    } else {
        std::cout << "NO PARTITION FOUND — counterexample!\n";
    }
    std::cout << "\n--- Timing ---\n";
    std::cout << "Table build (GPU): " << std::chrono::duration<double, std::milli>(t1 - t0).count() << " ms\n";
    std::cout << "Query (GPU)      : " << std::chrono::duration<double, std::milli>(t2 - t1).count() << " ms\n";
    std::cout << "Total            : " << std::chrono::duration<double, std::milli>(t2 - t0).count() << " ms\n";
    return p ? 0 : 1;
}
