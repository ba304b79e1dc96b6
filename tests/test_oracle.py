"""Pins the CPU oracle (oracle/liboracle.so) before anything is compared to it.

Sources of truth, in order: the reference's own fixtures (pi(10^k), primes
<= 30: src/test_sieve.cpp:31-40, src/test_bitset.cpp:17-50; the CLI run
goldbach 1e9 --p-small=1e5 -> 0 Phase-2 fallbacks: tests/validation_gpu.sh:48-54),
the reference's CPU sources compiled into oracle/_ref, literature values
(OEIS A025018/A025019 minimal-p records, Oliveira e Silva-Herzog-Pardi 2014),
the published partitions of RESULTS.md:114-121, and brute force.
"""
import os
import subprocess

import numpy as np
import pytest

from _helpers import REF_BIN, unpack

PI = {10: 4, 100: 25, 1000: 168, 10**4: 1229, 10**5: 9592, 10**6: 78498, 10**7: 664579, 10**8: 5761455}

# minimal-p records (n, p_min) up to 1e8: OEIS A025018 / A025019
RECORDS = [(4, 2), (6, 3), (12, 5), (30, 7), (98, 19), (220, 23), (308, 31), (556, 47), (992, 73), (2642, 103),
           (5372, 139), (7426, 173), (43532, 211), (54244, 233), (63274, 293), (113672, 313), (128168, 331),
           (194428, 359), (194470, 383), (413572, 389), (503222, 523), (1077422, 601), (3526958, 727),
           (3807404, 751), (10759922, 829), (24106882, 929), (27789878, 997), (37998938, 1039), (60119912, 1093)]

# published partitions (RESULTS.md:114-121): n = p + q with both prime
PUBLISHED = [(10**12, 194267), (10**13, 90059), (10**14, 145391), (10**15, 152639), (10**16, 56687),
             (10**17, 91967), (10**18, 14831), (10**19, 226283)]


def brute_pmin(n_max, is_prime):
    """p_min for every even n in [4, n_max] from a byte table, p ascending (cpu_goldbach.cpp:19-22 order)."""
    ns = np.arange(4, n_max + 1, 2, dtype=np.int64)
    pm = np.zeros(ns.shape, dtype=np.int64)
    pm[0] = 2  # 4 = 2 + 2
    todo = np.ones(ns.shape, dtype=bool)
    todo[0] = False
    for p in np.nonzero(is_prime)[0]:
        if p < 3:
            continue
        if not todo.any():
            break
        q = ns - p
        ok = todo & (q >= 2) & (q >= 0)
        hit = np.zeros_like(todo)
        hit[ok] = is_prime[q[ok]].astype(bool)
        pm[hit] = p
        todo &= ~hit
    return ns, pm


def test_pi_known_answers(orc):
    for n, pi in PI.items():
        assert len(orc.simple_sieve(n)) == pi
        assert int(orc.segmented_sieve(0, n).sum()) == pi
        if n >= 1000:
            words = orc.prime_bitset(n)
            nbits = (n - 3) // 2 + 1
            assert int(unpack(words, nbits).sum()) + 1 == pi


def test_primes_to_30(orc):
    words = orc.prime_bitset(30)
    bits = unpack(words, (30 - 3) // 2 + 1)
    primes = [2] + [2 * i + 3 for i in range(len(bits)) if bits[i]]
    assert primes == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]


def test_bitset_tail_bits_stay_set(orc):
    # constructor fills ~0 and the builder clears composites <= limit only
    for limit in (3, 4, 9, 30, 127, 128, 129, 130, 1001):
        nbits = (limit - 3) // 2 + 1
        w = orc.prime_bitset(limit)
        assert len(w) == (nbits + 63) // 64
        if nbits % 64:
            assert int(w[-1]) >> (nbits % 64) == (1 << (64 - nbits % 64)) - 1


@pytest.mark.parametrize("low,high", [(0, 10**6), (999_000, 1_001_000), (10**9, 10**9 + 100_000),
                                      (2**32 - 1000, 2**32 + 1000), (0, 1), (1, 2), (2, 2)])
def test_segmented_sieve_matches_reference(orc, ref, low, high):
    assert np.array_equal(orc.segmented_sieve(low, high), ref.segmented_sieve(low, high))


def test_simple_sieve_matches_reference(orc, ref):
    for lim in (0, 1, 2, 3, 100, 10**6 + 3):
        assert np.array_equal(orc.simple_sieve(lim), ref.simple_sieve(lim))


@pytest.mark.parametrize("limit", [0, 1, 2, 3, 4, 5, 9, 30, 64, 127, 128, 129, 1000, 10**6, 10**6 + 1, 10**7])
def test_prime_bitset_matches_reference(orc, ref, limit):
    # one OpenMP thread: the reference's builder races on shared words with more
    assert np.array_equal(orc.prime_bitset(limit), ref.prime_bitset(limit, threads=1))


def window_from_bytes(bytes_, q_low, q_high):
    odd = bytes_[::2][: (q_high - q_low) // 2 + 1]
    nbits = odd.shape[0]
    pad = (-nbits) % 64
    bits = np.concatenate([odd, np.ones(pad, dtype=np.uint8)])
    return np.packbits(bits, bitorder="little").view(np.uint64)


@pytest.mark.parametrize("q_low,q_high", [(3, 3), (3, 129), (3, 10**6 + 1), (101, 10001), (1_000_001, 1_065_535),
                                          (10**12 - 99_999, 10**12 + 200_001),
                                          (20_000_000_000_001, 20_000_000_200_001)])
def test_window_bits_match_reference_sieve(orc, ref, q_low, q_high):
    # the last case lies above the oracle's sieve table (sqrt > 2^22): its
    # Miller-Rabin path is checked against the reference's byte sieve
    got = orc.window_bits(q_low, q_high)
    want = window_from_bytes(ref.segmented_sieve(q_low, q_high), q_low, q_high)
    assert np.array_equal(got, want)


def trial_division(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    d = 3
    while d * d <= n:
        if n % d == 0:
            return False
        d += 2
    return True


def test_miller_rabin(orc, ref):
    table = ref.segmented_sieve(0, 200_000)
    got = np.array([orc.is_prime(n) for n in range(200_001)], dtype=np.uint8)
    assert np.array_equal(got, table)
    rng = np.random.default_rng(7)
    for n in rng.integers(2**31, 2**36, size=300):
        assert orc.is_prime(int(n)) == trial_division(int(n))
    # strong pseudoprimes to small base sets are composite
    for n in (2047, 1373653, 25326001, 3215031751, 2152302898747, 3474749660383, 341550071728321,
              3825123056546413051):
        assert not orc.is_prime(n)
    for p in (2**61 - 1, 2**64 - 59, 2**63 - 25, 18446744073709551557):
        assert orc.is_prime(p)
    assert not orc.is_prime(2**64 - 1) and not orc.is_prime(2**62 + 1)


# psi_k (OEIS A014233): the smallest odd composite that is a strong probable prime to
# each of the first k prime bases; below psi_k those k bases decide primality.  The
# device stops after mr_bases(n) bases on this table.
PSI = [(1, 2047), (2, 1373653), (3, 25326001), (4, 3215031751), (5, 2152302898747), (6, 3474749660383),
       (7, 341550071728321), (8, 341550071728321), (9, 3825123056546413051), (10, 3825123056546413051),
       (11, 3825123056546413051)]
PSI_FACTORS = {2047: (23, 89), 1373653: (829, 1657), 25326001: (2251, 11251), 3215031751: (151, 751, 28351),
               2152302898747: (6763, 10627, 29947), 3474749660383: (1303, 16927, 157543),
               341550071728321: (10670053, 32010157), 3825123056546413051: (149491, 747451, 34233211)}


def test_mr_base_count_thresholds(orc):
    for k, psi in PSI:
        f = PSI_FACTORS[psi]
        assert int(np.prod([np.uint64(x) for x in f], dtype=np.uint64)) == psi and all(trial_division(x) for x in f)
        assert orc.strong_first(psi, k), (k, psi)          # fools the first k bases ...
        assert not orc.is_prime(psi)                          # ... and not all twelve
    # where the k-base threshold moves, psi_k passes k bases but fails base k+1
    for k, psi in ((1, 2047), (2, 1373653), (3, 25326001), (4, 3215031751), (5, 2152302898747),
                   (6, 3474749660383), (8, 341550071728321), (11, 3825123056546413051)):
        assert not orc.strong_first(psi, k + 1), (k, psi)
    # minimality, by brute force, for the thresholds small enough to search
    assert orc.first_strong_pseudoprime(1, 2047) == 2047
    assert orc.first_strong_pseudoprime(2, 1373653) == 1373653


def test_published_partitions_are_prime(orc):
    for n, p in PUBLISHED:
        assert orc.is_prime(p) and orc.is_prime(n - p)


def test_min_partition_literature(orc):
    want = {10**12: 11, 10**13: 29, 10**14: 29, 10**15: 11, 10**16: 83, 10**17: 3, 10**18: 11, 10**19: 281,
            2**64 - 2: 277, 3325581707333960528: 9781}
    for n, p in want.items():
        assert orc.min_partition(n) == p
    for n, p in RECORDS:
        assert orc.min_partition(n) == p
    assert orc.min_partition(5) == 0 and orc.min_partition(2) == 0


def test_records_are_running_maxima(orc):
    # max p_min over [4, next_record - 2] equals the previous record, for records below 1e7
    for (n0, p0), (n1, _) in zip(RECORDS, RECORDS[1:]):
        if n1 > 10**7:
            break
        st = orc.verify_range(4, n1 - 2, 1 << 24, 10**6)
        assert (st["max_pmin"], st["argmax_n"]) == (p0, n0)


def test_verify_segment_vs_brute_force(orc, ref):
    n_max = 200_000
    table = ref.segmented_sieve(0, n_max + 10)
    ns, pm = brute_pmin(n_max, table)
    for A in (4, 6, 1000, 99_998):
        count = (n_max - A) // 2 + 1
        for p_small in (0, 2, 3, 5, 7, 100, 10**6):
            st, bits, pmin = orc.verify_segment(A, count, p_small)
            sub = pm[(A - 4) // 2:]
            assert np.array_equal(pmin.astype(np.int64), sub)
            ver = (sub <= p_small) & (sub > 0)
            assert np.array_equal(unpack(bits, count).astype(bool), ver)
            assert st["phase2"] == int((~ver).sum())
            assert st["failures"] == 0
            assert st["sum_pmin"] == int(sub.sum())


def test_sum_and_max_pmin_to_1e8(orc):
    st = orc.verify_range(4, 10**8, 1 << 24, 10**6)
    assert st["evens"] == 49_999_999
    assert (st["max_pmin"], st["argmax_n"]) == (1093, 60119912)
    assert st["sum_pmin"] == 1_511_603_116
    assert st["phase2"] == 0 and st["failures"] == 0


def test_reference_cli_fixture_1e9(orc):
    # validation_gpu.sh:48-54: goldbach 1e9 --seg-size=1e6 --p-small=1e5 -> 0 Phase-2 fallbacks
    st = orc.verify_range(4, 10**9, 10**6, 100_000)
    assert st["evens"] == 499_999_999 and st["phase2"] == 0 and st["failures"] == 0


def test_failure_path_hook(orc, ref):
    # with Phase 2 cut to p <= 7 the first n needing a larger p is the record n = 98
    table = ref.segmented_sieve(0, 1010)
    ns, pm = brute_pmin(1000, table)
    st, _, _ = orc.verify_segment(4, 499, 3, phase2_limit=7)
    assert st["first_fail_n"] == 98
    assert st["failures"] == int((pm > 7).sum())
    assert st["phase2"] == int((pm > 3).sum())
    st2 = orc.verify_range(4, 1000, 100, 3, phase2_limit=7)
    assert (st2["first_fail_n"], st2["failures"]) == (98, st["failures"])


def test_monotone_in_p_small(orc):
    prev = None
    for p_small in (0, 3, 7, 31, 100, 400, 10**6):
        st, bits, _ = orc.verify_segment(10**9, 1 << 16, p_small)
        cur = unpack(bits, 1 << 16).astype(bool)
        if prev is not None:
            assert not (prev & ~cur).any()
        prev = cur
    assert prev.all()


def test_cpu_goldbach_matches_reference_binary(orc):
    exe = os.path.join(REF_BIN, "cpu_goldbach")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built")
    out = subprocess.run([exe, "1000000"], capture_output=True, text=True, check=True).stdout
    assert "Even numbers checked : 499999" in out and "Failures found       : 0" in out
    assert "satisfy Goldbach. ✓" in out
    assert orc.cpu_goldbach(10**6) == (499_999, 0, 0)
    assert orc.cpu_goldbach(10**6 + 1) == (499_999, 0, 0)


def test_reference_known_answer_programs():
    for prog in ("test_sieve",):
        exe = os.path.join(REF_BIN, prog)
        if not os.path.exists(exe):
            pytest.skip("oracle/_ref not built")
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0 and "All tests passed." in r.stdout


# ---- partition counts c(n) (cpu_goldbach.cpp:29-36) ----

# Published numbers of Goldbach partitions of 10^k, k = 1..8 (OEIS A065577).
C_POW10 = (2, 6, 28, 127, 810, 5402, 38807, 291400)


def test_count_partitions_literature(orc):
    for k, want in enumerate(C_POW10, start=1):
        assert int(orc.count_partitions(10**k, 1)[0]) == want, k


def test_count_partitions_brute_force(orc):
    # tiny n by trial division, independent of every sieve here
    def isp(m):
        return m >= 2 and all(m % d for d in range(2, int(m**0.5) + 1))
    got = orc.count_partitions(4, 500)
    for i, c in enumerate(got):
        n = 4 + 2 * i
        assert int(c) == sum(1 for p in range(2, n // 2 + 1) if isp(p) and isp(n - p)), n
    # 4 = 2 + 2 and 6 = 3 + 3 are the only one-partition n below 12
    assert list(got[:4]) == [1, 1, 1, 2]


def test_count_partitions_matches_reference(orc, ref):
    for start, count in ((4, 3000), (999_000, 400), (10**7 - 200, 101)):
        assert np.array_equal(orc.count_partitions(start, count), ref.goldbach_count(start, count)), start


def test_count_partitions_positive_and_parity_sum(orc):
    # Every even n in [4, N] has a partition, and the sum of c(n) over them is the
    # number of prime pairs p <= q with p + q even and at most N (a double loop).
    N = 20_000
    got = orc.count_partitions(4, (N - 4) // 2 + 1)
    assert (got > 0).all()
    pr = [p for p in range(2, N) if all(p % d for d in range(2, int(p**0.5) + 1))]
    pairs = 0
    for i, p in enumerate(pr):
        for q in pr[i:]:
            if p + q > N:
                break
            if (p + q) % 2 == 0:
                pairs += 1
    assert int(got.sum()) == pairs


# Sinclair's seven bases, the device's Miller-Rabin set from psi_7 = 341550071728321 up
SINCLAIR = (2, 325, 9375, 28178, 450775, 9780504, 1795265022)


def test_wide_bases_agree_with_the_twelve_prime_bases(orc):
    """The device's wide base set against the oracle's twelve prime bases 2..37 on the
    strong pseudoprimes of the psi table, semiprimes of two primes near 2^32 (the hard
    composites for a strong test), the largest 64-bit prime, and random odd q up to 2^64."""
    def strong(n, a):
        d, s = n - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        x = pow(a, d, n)
        if x in (1, n - 1):
            return True
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                return True
        return False

    rng = np.random.default_rng(17)
    qs = [psi for _, psi in PSI if psi >= 341550071728321]
    qs += [int(x) | 1 for x in rng.integers(341550071728321, 2**63, size=4000, dtype=np.uint64)]
    qs += [int(x) * 2 + 1 for x in rng.integers(2**62, 2**63 - 1, size=4000, dtype=np.uint64)]
    near = [p for p in range(2**32 - 20000, 2**32) if trial_division(p)]
    qs += [near[i] * near[-1 - i] for i in range(200)]
    qs += [2**61 - 1, 2**64 - 59, (2**61 - 1) * 3 * 5 + 2]       # primes at the top, a near miss
    for q in qs:
        if q < 341550071728321 or q >= 2**64 or q % 2 == 0:
            continue
        assert all(strong(q, a) for a in SINCLAIR) == orc.is_prime(q), q
