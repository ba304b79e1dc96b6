"""The strong Lucas test as the device computes it (goldbach-gpu_amd/csrc/gb_device.hpp,
strong_lucas) against the textbook form (Baillie & Wagstaff 1980, Selfridge's method A).

The device runs its ladder on W_k = V_2k / Q^k, the Q = 1 sequence with parameter 1/Q - 2, takes
D from a table of Jacobi symbols over residues of n (every Selfridge D is 1 mod 4, so
(D/n) = (n mod |D| / |D|)) and reads U_d = 0 and V_(d 2^r) = 0 off the W values.  Both forms
are restated here in plain modular arithmetic; they must agree on every odd n, on the strong
Lucas pseudoprimes (OEIS A217255) and on random 64-bit n.  The Baillie-PSW answer (base-2
strong test and strong Lucas) must equal the oracle's on random 64-bit n.  The device form
is the product path for q above 2^44 in the tiles (goldbach.cu:119-158 is the reference's
twelve-base test those tiles must agree with)."""
import math
import random

import numpy as np

# OEIS A217255: strong Lucas pseudoprimes (Selfridge's parameters)
STRONG_LUCAS_PSP = [5459, 5777, 10877, 16109, 18971, 22499, 24569, 25199, 40309, 58519, 75077, 97439]
# residue groups of the device's D table: 3^3 5^2 7 11 13 17 19, 23 29 31 37 41, 43 47
GROUPS = [218243025, 31367009, 2021]


def jacobi(a: int, m: int) -> int:
    a %= m
    t = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if m % 8 in (3, 5):
                t = -t
        a, m = m, a
        if a % 4 == 3 and m % 4 == 3:
            t = -t
        a %= m
    return t if m == 1 else 0


def kronecker_d(D: int, n: int) -> int:
    """(D/n) for odd n > 0 by the general rule: reciprocity for |D|, then (-1/n)."""
    ad = abs(D)
    j = jacobi(n % ad, ad)
    if (ad - 1) // 2 % 2 == 1 and (n - 1) // 2 % 2 == 1:
        j = -j
    if D < 0 and n % 4 == 3:
        j = -j
    return j


def selfridge(n: int):
    """First D of 5, -7, 9, ... with (D/n) = -1; None when some (D/n) = 0 comes first (or n is a square)."""
    r = math.isqrt(n)
    if r * r == n:
        return None
    D = 5
    while True:
        j = kronecker_d(D, n)
        if j == -1:
            return D
        if j == 0 and abs(D) != n:
            return None
        D = -D - 2 if D > 0 else -D + 2


def lucas_uv(n: int) -> bool:
    """Textbook strong Lucas test: U_d = 0 or V_(d 2^r) = 0 for some 0 <= r < s, n + 1 = d 2^s."""
    D = selfridge(n)
    if D is None:
        return False
    Q = (1 - D) // 4
    d, s = n + 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    half = lambda x: x // 2 if x % 2 == 0 else (x + n) // 2
    U, V, Qk = 1, 1, Q % n
    for b in bin(d)[3:]:
        U, V, Qk = U * V % n, (V * V - 2 * Qk) % n, Qk * Qk % n
        if b == "1":
            U, V = half((U + V) % n), half((D * U + V) % n)
            Qk = Qk * Q % n
    if U == 0 or V == 0:
        return True
    for _ in range(s - 1):
        V = (V * V - 2 * Qk) % n
        Qk = Qk * Qk % n
        if V == 0:
            return True
    return False


def table_d(n: int):
    """The device's D search: the 22 candidates |D| <= 47 from residues of n, then the plain search."""
    res = [n % g for g in GROUPS]
    D = 5
    while abs(D) <= 47:
        m = abs(D)
        a = res[[g for g in range(3) if GROUPS[g] % m == 0][0]] % m
        j = jacobi(a, m)                       # = (D/n) because D = 1 mod 4
        if j == 0:
            return None
        if j == -1:
            return D
        D = -D - 2 if D > 0 else -D + 2
    r = math.isqrt(n)
    if r * r == n:
        return None
    while True:
        j = jacobi(n % abs(D), abs(D))
        if j == -1:
            return D
        if j == 0:
            return None
        D = -D - 2 if D > 0 else -D + 2


def lucas_w(n: int) -> bool:
    """The device's form: ladder on W_k = V_2k / Q^k (parameter P' = 1/Q - 2, Q' = 1)."""
    D = table_d(n)
    if D is None:
        return False
    Q = (1 - D) // 4
    if math.gcd(abs(Q), n) != 1:
        return False
    pp = (pow(Q, -1, n) - 2) % n
    d = n + 1
    s = (d & -d).bit_length() - 1
    d >>= s
    m = d >> 1
    w0, w1 = 2, pp                              # W_0, W_1
    top = -1
    if m:
        w0, w1 = pp, (pp * pp - 2) % n          # top bit of m: W_1, W_2
        top = m.bit_length() - 2
    for b in range(top, -1, -1):
        x = (w0 * w1 - pp) % n                  # W_2k+1
        if (m >> b) & 1:
            w0, w1 = x, (w1 * w1 - 2) % n
        else:
            w0, w1 = (w0 * w0 - 2) % n, x
    if w0 == w1 or (w0 + w1) % n == 0:          # U_d = 0, V_d = 0
        return True
    w = (w0 * w1 - pp) % n                      # W_d
    for _ in range(1, s):
        if w == 0:                              # V_(d 2^i) = 0
            return True
        w = (w * w - 2) % n
    return False


def sprp2(n: int) -> bool:
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(2, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def test_reciprocity_shortcut_for_selfridge_d():
    rng = random.Random(9)
    Ds, D = [], 5
    while abs(D) < 500:
        Ds.append(D)
        D = -D - 2 if D > 0 else -D + 2
    assert all(D % 4 == 1 for D in Ds)
    for _ in range(400):
        n = rng.randrange(3, 2**64) | 1
        for D in Ds:
            assert kronecker_d(D, n) == jacobi(n % abs(D), abs(D)), (D, n)


def test_pseudoprimes_pass_both_forms():
    for n in STRONG_LUCAS_PSP:
        assert lucas_uv(n) and lucas_w(n) and not sprp2(n), n


def test_forms_agree_on_every_odd_n():
    bad = [n for n in range(101, 60001, 2) if lucas_uv(n) != lucas_w(n)]
    assert bad == []
    # and the brute-force count of strong Lucas pseudoprimes below 60,000 is the OEIS one
    sieve = np.ones(60001, dtype=bool)
    sieve[:2] = False
    for i in range(2, 245):
        if sieve[i]:
            sieve[i * i::i] = False
    psp = [n for n in range(101, 60001, 2) if not sieve[n] and lucas_w(n)]
    assert psp == [x for x in STRONG_LUCAS_PSP if x <= 60000]


def test_forms_agree_on_64_bit_n():
    rng = random.Random(7)
    cases = [rng.randrange(2**33, 2**64) | 1 for _ in range(6000)]
    cases += [2**31 - 1, 2**61 - 1, 2**61 - 3, 2**62 - 1, 2**63 - 25, 2**64 - 59,
              4194319**2, 2147483647**2, 341550071728321, 3825123056546413051,
              8608456956238879741, 11682905869181336791]        # the last two need |D| > 47
    for n in cases:
        assert lucas_uv(n) == lucas_w(n), n


def test_baillie_psw_matches_the_oracle(orc):
    rng = random.Random(11)
    ns = np.array([rng.randrange(2**44, 2**64) | 1 for _ in range(3000)]
                  + [341550071728321, 3825123056546413051, 2**61 - 1], dtype=np.uint64)
    want = [orc.is_prime(int(n)) for n in ns]
    for n, w in zip(ns.tolist(), want):
        assert (sprp2(n) and lucas_w(n)) == bool(w), n
