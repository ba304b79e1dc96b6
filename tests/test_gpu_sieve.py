"""Device sieve parity: gb_sieve_window against the oracle's window bits
(reference kernel contract: src/goldbach.cu:198-259, tail bits 1) and the
PrimeBitset layout (include/prime_bitset.hpp of the reference)."""
import numpy as np
import pytest

from _helpers import unpack

pytestmark = pytest.mark.gpu

WINDOWS = [
    (3, 3), (3, 5), (3, 127), (3, 129), (3, 131), (3, 1001), (61, 63), (67, 4491), (3, 10**6 + 1),
    (4489, 4489 + 2 * 64 * 1000 - 2),               # starts at 67^2
    (10**9 + 1, 10**9 + 200_001),
    (999_999_999_001, 1_000_000_077_777),           # around 1e12 (sieve with p <= 1e6)
    (17_592_186_044_417, 17_592_186_244_417),       # just above 2^44: pre-sieve + Miller-Rabin
    (17_592_177_655_809, 17_592_186_244_417),       # straddles 2^44: some tiles sieve, some test
    (10**18 + 1, 10**18 + 100_001),
    (2**64 - 2**21 + 1, 2**64 - 1),                 # top of the 64-bit range
    (2**61 - 1 - 100_000, 2**61 - 1 + 100_000),     # Mersenne prime: n + 1 = 2^61, the Lucas ladder is empty
    (341550071728321 - 100_000, 341550071728321 + 100_000),          # strong pseudoprime to bases 2..19
    (3825123056546413051 - 40_000, 3825123056546413051 + 40_000),    # ... to bases 2..31
]

# Base-2 strong pseudoprimes above 2^44: 10670053 * 32010157 and 4195117 * 8390233 have
# every factor above the 2^22 pre-sieve, so they reach the tiles' second test (the strong
# Lucas test of Baillie-PSW); the pre-sieve removes 262957 * 525913 * 788869 (a
# (6k+1)(12k+1)(18k+1) Carmichael number).
SPSP = [341550071728321, 35198009092261, 109094669922527929]
# squares of primes above the pre-sieve: no Lucas parameter D exists for them
SQUARES = [4194319**2, 2147483647**2]


@pytest.mark.parametrize("q", SPSP + SQUARES)
def test_strong_pseudoprime_is_cleared(ctx, orc, q):
    bits = ctx.sieve_window(q - 2 * 64, q + 2 * 64)
    assert unpack(bits, 129)[64] == 0
    assert orc.strong_first(q, 1) == (q in SPSP) and not orc.is_prime(q)
    # and the window around it equals the oracle's
    assert np.array_equal(ctx.sieve_window(q - 20_000, q + 20_000), orc.window_bits(q - 20_000, q + 20_000))


# Primes above 2^44 by the Selfridge parameter D their strong Lucas test takes, one for each
# odd part c of |Q| = |1 - D|/4 > 1 in the device's D table (built by the Chinese remainder
# theorem: q = 1 modulo every odd prime below |D|, a non-residue modulo |D|): D = 29 (c = 7),
# 37 (c = 9), 41 (c = 5), -43 (c = 11), -47 (c = 3); and primes 1 + k * 47#/2 (k = 28, 38),
# 1 modulo every odd prime up to 47, whose search runs past the table (D = 53 and -59).
LATE_D = [1125910450544881, 1133367329724631, 5276144813849911, 53852296644316171, 3480014514224227981,
          8608456956238879741, 11682905869181336791]


@pytest.mark.parametrize("q", LATE_D)
def test_lucas_parameter_cases(ctx, orc, q):
    assert orc.is_prime(q)
    bits = ctx.sieve_window(q - 2 * 64, q + 2 * 64)
    assert unpack(bits, 129)[64] == 1
    assert np.array_equal(ctx.sieve_window(q - 40_000, q + 40_000), orc.window_bits(q - 40_000, q + 40_000))


@pytest.mark.parametrize("q_low,q_high", WINDOWS)
def test_window_matches_oracle(ctx, orc, q_low, q_high):
    got = ctx.sieve_window(q_low, q_high)
    want = orc.window_bits(q_low, q_high)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("limit", [3, 4, 5, 9, 30, 127, 128, 129, 130, 10**6, 10**6 + 1, 10**8])
def test_prime_bitset_layout(ctx, orc, limit):
    q_high = limit if limit & 1 else limit - 1
    got = ctx.sieve_window(3, q_high)
    assert np.array_equal(got, orc.prime_bitset(limit))


def test_pi_1e9_full_window(ctx):
    # chunked through the segment buffer; pi(1e9) = 50,847,534 (2 is not in the odd set)
    bits = ctx.sieve_window(3, 999_999_999)
    nbits = (999_999_999 - 3) // 2 + 1
    assert int(unpack(bits, nbits).sum()) == 50_847_534 - 1


@pytest.mark.parametrize("q_low,q_high", [(2, 11), (1, 11), (5, 4), (3, 10), (4, 11)])
def test_window_rejects_bad_bounds(ctx, gb, q_low, q_high):
    with pytest.raises(gb.GoldbachError) as e:
        ctx.sieve_window(q_low, q_high)
    assert e.value.status == gb.GB_EINVAL
