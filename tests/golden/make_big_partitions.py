#!/usr/bin/env python3
"""Writes big_partitions.json: even n above 2^64 with the smallest prime p <= 10^7
whose n - p is a strong probable prime to the bases 2..97, from oracle/bigcheck.py,
each answer re-checked with the image's GMP mpz_probab_prime_p(q, 25) (the reference's
test, big_check.cpp:107): GMP accepts n - p and rejects n - p' for every prime p' < p.

The n: the published big_check rows (RESULTS.md:141-148: 10^50, 10^100, 10^1000;
10^200 is listed with p = 113, but 13 divides 10^200 - 113), powers of two where
n - p loses its top word, and seeded random n of 3 to 130 32-bit words (the sizes
where the device's column and carry blocks change shape)."""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _helpers import Gmp, load_bigcheck  # noqa: E402


def cases():
    out = [("10^50", 10**50), ("10^100", 10**100), ("10^200", 10**200), ("10^1000", 10**1000),
           ("2^64", 2**64), ("2^64+2", 2**64 + 2), ("2^96", 2**96), ("2^128-2", 2**128 - 2),
           ("2^4128", 2**4128), ("2^4096+2^4064*7", 2**4096 + 7 * 2**4064)]
    rng = random.Random(20261204)
    for words in (3, 4, 5, 8, 16, 17, 33, 63, 64, 65, 127, 128, 129, 130):
        n = rng.getrandbits(32 * words) | (1 << (32 * words - 1))
        out.append((f"random {words} words", n & ~1))
    return out


def main():
    bc = load_bigcheck()
    gmp = Gmp()
    primes = bc.primes_upto(10**7)
    rows = []
    for name, n in cases():
        p = bc.min_partition(n, 25, 10**7, primes)
        assert p > 0, name
        assert gmp.probab_prime(n - p, 25) > 0, name
        for r in primes:
            if r >= p:
                break
            assert gmp.probab_prime(n - r, 25) == 0, (name, r)
        rows.append({"name": name, "n": str(n), "p": p})
        print(name, p, flush=True)
    with open(os.path.join(HERE, "big_partitions.json"), "w") as f:
        json.dump({"gmp": gmp.version, "rounds": 25, "p_max": 10**7, "cases": rows}, f, indent=1)


if __name__ == "__main__":
    main()
