"""gb_verify_segment_noalloc: the allocation-free variant of gb_verify_segment (static device
scratch, segments of at most 2^20 evens) gives the same bitmaps, p_min arrays and statistics as
gb_verify_segment, call after call, larger segments before smaller ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin", "argmax_n")

# largest first, so a later call meets what the earlier one left in the shared scratch
CASES = [
    (10**12 - 2 * (1 << 20), 1 << 20, 10**6),        # the scratch's largest segment, near 1e12
    (10**15, 20_000, 10**6),                          # tiles decided by pre-sieve + Baillie-PSW
    (3325581707333960528 - 600, 700, 10**6),          # p_min 9781: the slow path
    (4, 1000, 7),                                     # small P_SMALL: Phase-2 counts
    (2**64 - 2 - 2 * 4095, 4096, 10**6),              # the top of the 64-bit range
]


def same(a, b):
    for k in KEYS:
        assert a.stats[k] == b.stats[k], k
    if a.verified is not None:
        assert np.array_equal(a.verified, b.verified)
    if a.pmin is not None:
        assert np.array_equal(a.pmin, b.pmin)


def test_noalloc_equals_verify_segment(gb):
    with gb.Context(0) as c:
        for A, count, p_small in CASES:
            got = c.verify_segment_noalloc(A, count, p_small, want_pmin_array=True)
            want = c.verify_segment(A, count, p_small, want_pmin_array=True)
            same(got, want)
            # the unrolled marking path (no per-n outputs) and a stats-only call
            same(c.verify_segment_noalloc(A, count, p_small), c.verify_segment(A, count, p_small))
            s = c.verify_segment_noalloc(A, count, p_small, want_verified=False)
            assert all(s.stats[k] == want.stats[k] for k in ("evens", "phase2", "failures", "first_fail_n"))


def test_noalloc_two_calls_shrinking(gb, orc):
    with gb.Context(0) as c:
        big = c.verify_segment_noalloc(4, 1 << 20, 10**6, want_pmin_array=True)
        small = c.verify_segment_noalloc(10**9, 5000, 10**6, want_pmin_array=True)
        for (A, count), r in (((4, 1 << 20), big), ((10**9, 5000), small)):
            st, bits, pmin = orc.verify_segment(A, count, 10**6)
            assert np.array_equal(r.verified, bits) and np.array_equal(r.pmin, pmin)
            for k in KEYS:
                assert r.stats[k] == st[k], k
        assert small.stats["evens"] == 5000


def test_noalloc_rejects_more_than_its_scratch_holds(gb):
    with gb.Context(0) as c:
        with pytest.raises(gb.GoldbachError) as e:
            c.verify_segment_noalloc(4, (1 << 20) + 1, 10**6)
        assert e.value.status == gb.GB_ERANGE and "2^20" in e.value.message
        with pytest.raises(gb.GoldbachError) as e:
            c.verify_segment_noalloc(5, 10, 10**6)
        assert e.value.status == gb.GB_EINVAL
        # the context still answers after the refusals
        assert c.verify_segment_noalloc(4, 64, 10**6).stats["evens"] == 64
