"""Range statistics fixtures (tests/golden/range_sums.json, written by the oracle through
tests/golden/make_range_sums.py) and the published minimal-partition records they contain.

The largest p_min of each range from 4 (and of [1e10, 1e10 + 2^31]) is a record of T. Oliveira
e Silva's table of minimal Goldbach partitions (OEIS A025018 for n, A025019 for p): 1789 at
721,013,438; 1877 at 7,473,202,036; 1879 at 11,001,080,372; 3457 at 599,533,546,358, the
largest p_min below 1e12.  The oracle's per-n scan (cpu_goldbach.cpp:19-22 order) must give
exactly those, and [4, 1e9] must be reproducible from the oracle here.  [1e12 - 2^31, 1e12]
is no record range: its maximum is only checked to be the stated minimal partition."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RECORDS = [(721_013_438, 1789), (7_473_202_036, 1877), (11_001_080_372, 1879), (599_533_546_358, 3457)]


def fixtures():
    with open(os.path.join(HERE, "golden", "range_sums.json")) as f:
        return json.load(f)


def test_fixture_shape():
    assert {(r["start"], r["limit"]) for r in fixtures()} >= {(4, 10**9), (4, 10**10), (4, 10**12),
                                                             (10**12 - 2**31, 10**12)}
    for r in fixtures():
        assert r["evens"] == (r["limit"] - r["start"]) // 2 + 1
        assert r["phase2"] == 0 and r["failures"] == 0 and r["first_fail_n"] == 0
        assert r["start"] <= r["argmax_n"] <= r["limit"]


def test_records_are_the_ranges_maxima(orc):
    # a range from 4 ends its running maximum at a published record; every range's argmax has
    # exactly the stated minimal partition
    rec = dict(RECORDS)
    for r in fixtures():
        n, p = r["argmax_n"], r["max_pmin"]
        if r["start"] == 4:
            assert rec.get(n) == p, (r["start"], r["limit"])
        assert orc.min_partition(n) == p
        assert orc.is_prime(p) and orc.is_prime(n - p)
    assert {(r["argmax_n"], r["max_pmin"]) for r in fixtures()} >= set(RECORDS)


@pytest.mark.parametrize("idx", [0])
def test_fixture_reproduces(orc, idx):
    r = fixtures()[idx]
    st = orc.verify_range(r["start"], r["limit"], 1 << 27, r["p_small"])
    for k in ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin", "argmax_n"):
        assert st[k] == r[k], k
