"""gb_big_min_partition on the device against the committed fixture (oracle/
bigcheck.py answers, each confirmed by GMP's mpz_probab_prime_p, the test of
big_check.cpp:107): published rows (RESULTS.md:141-148), n whose n - p loses its
top word, random n of 3..130 words; below 2^64 against the 64-bit path; rounds
and p_max limits; argument errors; and the big_check CLI contract
(tests/validation.sh:43-49, 97-107)."""
import json
import os
import subprocess

import numpy as np
import pytest

from _helpers import PKG_DIR

pytestmark = pytest.mark.gpu
BIG = os.path.join(PKG_DIR, "bin", "big_check")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "big_partitions.json")


def cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_fixture(ctx, case):
    n = int(case["n"])
    p, st = ctx.big_min_partition(n)
    assert p == case["p"], (case["name"], p)
    assert st["words"] == (n.bit_length() + 31) // 32
    assert st["mr_tasks"] >= st["candidates"] and st["candidates"] <= st["primes_scanned"]


def test_below_2_64(ctx, orc):
    rng = np.random.default_rng(5)
    ns = [4, 6, 10**12, 2**64 - 2] + [int(x) * 2 for x in rng.integers(2, 2**62, size=30)]
    ref = orc.min_partition_batch(ns)
    for n, p in zip(ns, ref):
        assert ctx.big_min_partition(n)[0] == int(p), n


def test_rounds_and_p_max(ctx):
    n = 10**300
    p25, _ = ctx.big_min_partition(n)
    p1, _ = ctx.big_min_partition(n, rounds=1)
    assert p1 == p25                       # no base-2 strong pseudoprime in the way here
    assert ctx.big_min_partition(n, p_max=p25)[0] == p25
    assert ctx.big_min_partition(n, p_max=p25 - 1)[0] == 0


def test_argument_errors(ctx, gb):
    for n, kw in ((2**70 + 1, {}), (2, {}), (0, {}), (2**(32 * 2304), {}), (2**70, {"rounds": 26}),
                  (2**70, {"p_max": 2**32})):
        with pytest.raises(gb.GoldbachError):
            ctx.big_min_partition(n, **kw)


def run(args):
    r = subprocess.run([BIG] + args, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout + r.stderr


def test_cli_contract():
    rc, out = run(["1000000000000"])
    assert rc == 0 and "Goldbach holds. ✓" in out and "= 11 + 999999999989" in out
    rc, out = run(["1" + "0" * 50])
    assert rc == 0 and "Goldbach holds. ✓" in out and "= 383 + " in out
    rc, out = run(["1" + "0" * 1000])
    assert rc == 0 and "p = 26981" in out and "q has 1000 digits" in out
    rc, out = run(["1000000000000000000000000000001"])
    assert rc == 1 and "Error: Goldbach's conjecture applies to even" in out
    rc, out = run(["abc"])
    assert rc == 1 and "Error: Input must be a positive integer" in out
    rc, out = run(["2"])
    assert "Error: invalid number or not even / >= 4" in out
