// big_check_main.cpp -- the `big_check` command line, drop-in for the
// reference's arbitrary-precision checker (src/big_check.cpp:35-175): same
// argument checks, messages and result lines, with the primality tests of
// n - p run on the device by gb_big_min_partition (strong tests to the bases
// 2..97 in place of GMP's mpz_probab_prime_p(q, 25)).  The p printed is the
// smallest prime <= 10^7 whose n - p passes; the reference's is the first a
// thread reports (RESULTS.md:133-135).
//
//   big_check <N> [--rounds=K] [--p-max=P]
#include <cctype>
#include <chrono>
#include <cstdint>
#include <iostream>
#include <string>
#include <vector>

#include "goldbach_gpu.h"

int main(int argc, char** argv) {
    if (argc < 2 || std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help") {
        std::cout << "Arbitrary Precision Goldbach Checker (MI355X)\n";
        std::cout << "Usage:   " << argv[0] << " <N> [--rounds=K] [--p-max=P]\n";
        std::cout << "Example: " << argv[0] << " 1000000000000000000000000000000\n";
        return 0;
    }
    const std::string n_str = argv[1];
    uint32_t rounds = 25;
    uint64_t p_max = 10'000'000ull;
    try {
        for (int i = 2; i < argc; ++i) {
            const std::string a = argv[i];
            if (a.rfind("--rounds=", 0) == 0) rounds = (uint32_t)std::stoul(a.substr(9));
            else if (a.rfind("--p-max=", 0) == 0) p_max = std::stoull(a.substr(8));
        }
    } catch (...) {
        std::cerr << "Error: Invalid numeric argument.\n";
        return 1;
    }
    for (char ch : n_str) {
        if (!std::isdigit((unsigned char)ch)) {
            std::cerr << "Error: Input must be a positive integer.\n";
            return 1;
        }
    }
    const int last_digit = n_str.back() - '0';
    if (last_digit % 2 != 0) {
        std::cerr << "Error: Goldbach's conjecture applies to even integers >= 4.\n";
        std::cerr << "Input ends in '" << last_digit << "' (odd).\n";
        return 1;
    }

    if (n_str.size() <= 100) std::cout << "Checking Goldbach for n = " << n_str << "\n";
    else std::cout << "Checking Goldbach for n with " << n_str.size() << " digits\n";
    std::cout << "  (" << n_str.size() << " digits)\n\n";

    std::vector<uint32_t> n(n_str.size() / 9 + 2);
    uint32_t nw = 0;
    if (gb_big_from_decimal(n_str.c_str(), n.data(), (uint32_t)n.size(), &nw) != GB_OK ||
        (nw == 1 && n[0] < 4)) {
        std::cerr << "Error: invalid number or not even / >= 4\n";
        return 0;                        // the reference returns normally here (big_check.cpp:50-54)
    }
    const auto t0 = std::chrono::steady_clock::now();
    gb_ctx* ctx = nullptr;
    if (gb_open(0, nullptr, &ctx) != GB_OK) {
        std::cerr << "Error: " << gb_last_error(nullptr) << "\n";
        return 1;
    }
    char name[256] = {0};
    int cus = 0;
    gb_device_info(0, name, sizeof name, nullptr, &cus);
    std::cout << "Testing n - p for primes p up to " << p_max << " on GPU 0 (" << name << ", " << cus
              << " CUs), " << rounds << " strong-test bases...\n\n";
    uint64_t p = 0;
    gb_big_stats st{};
    const int rc = gb_big_min_partition(ctx, n.data(), nw, rounds, p_max, &p, &st);
    if (rc != GB_OK) {
        std::cerr << "Error: " << gb_last_error(ctx) << "\n";
        gb_close(ctx);
        return 1;
    }
    gb_close(ctx);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    std::cout << "\n--- Result ---\n";
    if (p > 0) {
        // q = n - p, then its digits
        std::vector<uint32_t> q(n.begin(), n.begin() + nw);
        uint64_t sub = p, borrow = 0;
        for (size_t j = 0; j < q.size(); ++j) {
            const uint64_t d = (uint64_t)q[j] - (sub & 0xFFFFFFFFull) - borrow;
            q[j] = (uint32_t)d;
            borrow = (d >> 63) & 1;
            sub >>= 32;
        }
        std::vector<char> digits(q.size() * 10 + 2);
        gb_big_to_decimal(q.data(), (uint32_t)q.size(), digits.data(), digits.size());
        const std::string q_str(digits.data());
        if (n_str.size() <= 100) {
            std::cout << n_str << "\n  = " << p << " + " << q_str << "\n";
        } else {
            std::cout << "n has " << n_str.size() << " digits\n";
            std::cout << "p = " << p << "\n";
            std::cout << "q has " << q_str.size() << " digits\n";
        }
        std::cout << "Goldbach holds. ✓\n";
    } else {
        std::cout << "NO PARTITION FOUND -- counterexample!\n";
    }
    std::cout << "\n--- Timing ---\n";
    std::cout << "Primes scanned : " << st.primes_scanned << " (" << st.candidates
              << " with n - p free of factors below 2^20)\n";
    std::cout << "Strong tests   : " << st.mr_tasks << " in " << st.batches << " launch(es)\n";
    std::cout << "Device time    : " << st.kernel_ms << " ms\n";
    std::cout << "Total time   : " << ms / 1000.0 << " seconds\n";
    return 0;
}
