"""The arbitrary-size single-n search (big_check.cpp:35-144) without a device:
the Python restatement in oracle/bigcheck.py against the image's GMP (the
reference's dependency, mpz_probab_prime_p(q, 25) at big_check.cpp:107), the
reference's published rows (RESULTS.md:141-148), the 64-bit oracle below 2^64,
the committed fixture, and the library's decimal conversion against Python's."""
import json
import os
import random
import sys

import numpy as np
import pytest

from _helpers import Gmp, load_bigcheck

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "big_partitions.json")
if hasattr(sys, "set_int_max_str_digits"):
    sys.set_int_max_str_digits(0)


@pytest.fixture(scope="module")
def bc():
    return load_bigcheck()


@pytest.fixture(scope="module")
def gmp():
    try:
        return Gmp()
    except OSError:
        pytest.skip("libgmp.so.10 not present")


def test_published_rows(bc):
    # 10^50 -> 383 and 10^100 -> 797 as published; the single-thread 10^1000 row (p = 26,981) is in the fixture
    primes = bc.primes_upto(10**4)
    assert bc.min_partition(10**50, primes=primes) == 383
    assert bc.min_partition(10**100, primes=primes) == 797
    # the published 10^200 -> 113 is not a partition: 13 divides 10^200 - 113
    assert (10**200 - 113) % 13 == 0
    assert bc.min_partition(10**200, primes=primes) == 7103


def test_strong_test_against_gmp(bc, gmp):
    rng = random.Random(7)
    qs = [rng.getrandbits(b) | 1 for b in (40, 64, 65, 100, 200, 520, 1100) for _ in range(30)]
    # strong pseudoprimes to several bases and Carmichael numbers: GMP and the 25 bases both reject
    qs += [2047, 1373653, 25326001, 3215031751, 2152302898747, 3474749660383, 341550071728321,
           3825123056546413051, 318665857834031151167461, 3317044064679887385961981, 561, 41041, 825265]
    for q in qs:
        assert bc.probable_prime(q) == (gmp.probab_prime(q, 25) > 0), q


def test_below_2_64_matches_the_64_bit_oracle(bc, orc):
    rng = random.Random(3)
    ns = [4, 6, 8, 98, 10**12, 2**64 - 2, 3325581707333960528] + [rng.randrange(4, 2**64, 2) for _ in range(40)]
    got = orc.min_partition_batch(ns)
    primes = bc.primes_upto(10**5)
    for n, p in zip(ns, got):
        assert bc.min_partition(n, primes=primes) == int(p), n


def test_fixture(bc, gmp):
    with open(GOLDEN) as f:
        fx = json.load(f)
    assert fx["rounds"] == 25 and fx["p_max"] == 10**7 and len(fx["cases"]) >= 20
    primes = bc.primes_upto(10**5)
    for case in fx["cases"]:
        n, p = int(case["n"]), case["p"]
        assert gmp.probab_prime(n - p, 25) > 0, case["name"]
        if n.bit_length() <= 32 * 17:            # recompute the cheap ones
            assert bc.min_partition(n, primes=primes) == p, case["name"]


def test_decimal_conversion(gb):
    rng = random.Random(11)
    for digits in (1, 2, 9, 10, 18, 19, 20, 100, 1001, 10001):
        for _ in range(3):
            s = str(rng.randrange(10 ** (digits - 1), 10**digits)) if digits > 1 else str(rng.randrange(10))
            w = gb.from_decimal(s)
            v = int(s)
            assert int.from_bytes(w.astype("<u4").tobytes(), "little") == v
            assert len(w) == max(1, (v.bit_length() + 31) // 32)
            assert gb.to_decimal(w) == s
    assert gb.to_decimal(gb.from_decimal("000123")) == "123"
    assert gb.to_decimal(gb.from_decimal("0")) == "0"
    assert gb.to_decimal(np.array([0, 0, 0], dtype=np.uint32)) == "0"
    for bad in ("", "12a", "-4", " 4"):
        with pytest.raises(gb.GoldbachError) as e:
            gb.from_decimal(bad)
        assert e.value.status == gb.GB_EINVAL
    with pytest.raises(gb.GoldbachError) as e:
        gb.from_decimal("9" * 30, cap=3)
    assert e.value.status == gb.GB_ERANGE
    assert np.array_equal(gb.big_words(2**64 + 5), np.array([5, 0, 1], dtype=np.uint32))
