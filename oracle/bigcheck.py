"""CPU restatement of big_check's search for arbitrarily large even n -- TEST
INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's CPU baseline use it as the
checker; the product never imports it).

Reference: big_check.cpp:35-144.  The reference scans the primes p <= 10^7
(:58-61) in batches of 1,000 over OpenMP threads (:72-118) and accepts the
first p whose q = n - p passes GMP's mpz_probab_prime_p(q, 25) (:107; GMP
6.3.0 per RESULTS.md:15: Baillie-PSW plus one Miller-Rabin round with a
pseudo-random base).  Its p is the first a thread reports, not always the
smallest of the batch.

Restated here: the smallest prime p <= p_max such that q = n - p > 1 is a
strong probable prime to each of the first `rounds` primes 2, 3, 5, ... 97
(Miller-Rabin with fixed bases; Python's three-argument pow does the
modular exponentiation).  For every q the tests meet, GMP's own
mpz_probab_prime_p (the image's libgmp) gives the same verdict, and below
2^64 the bases 2..37 decide primality exactly.
"""
from __future__ import annotations

BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97)


def strong_probable_prime(q: int, a: int) -> bool:
    """Miller-Rabin to base a: q - 1 = 2^s d, a^d = 1 or a^(2^r d) = -1 (mod q) for some r < s."""
    if q < 2:
        return False
    if q % 2 == 0:
        return q == 2
    if q % a == 0:
        return q == a
    d, s = q - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(a, d, q)
    if x == 1 or x == q - 1:
        return True
    for _ in range(s - 1):
        x = x * x % q
        if x == q - 1:
            return True
        if x == 1:
            return False
    return False


def probable_prime(q: int, rounds: int = 25) -> bool:
    return q >= 2 and all(strong_probable_prime(q, a) for a in BASES[:rounds])


def primes_upto(limit: int) -> list:
    if limit < 2:
        return []
    sieve = bytearray([1]) * (limit + 1)
    sieve[0] = sieve[1] = 0
    for i in range(2, int(limit ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(sieve[i * i::i]))
    return [i for i in range(limit + 1) if sieve[i]]


_SMALL = primes_upto(1 << 12)


def min_partition(n: int, rounds: int = 25, p_max: int = 10**7, primes: list | None = None) -> int:
    """Smallest prime p <= min(p_max, n/2) with n - p a probable prime (0 if none); big_check.cpp:83-118."""
    if n < 4 or n % 2:
        raise ValueError("n must be even and >= 4")
    for p in primes if primes is not None else primes_upto(p_max):
        if p > p_max or 2 * p > n:
            break
        q = n - p
        # trial division first: a q with a small factor other than itself is composite
        if any(q % l == 0 and q != l for l in _SMALL):
            continue
        if probable_prime(q, rounds):
            return p
    return 0
