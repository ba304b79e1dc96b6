"""The plain form's kernel time.  Plain segments end without an event between them: each
segment's reduction writes the device wall clock at the segment's start and end and then a
sequence number into pinned memory, and the host waits on that word.  So the kernel time a
plain call reports comes from the device clock (the union of its segments' intervals, since two
segments run side by side on two streams): it must be positive, no longer than the call,
identical in its statistics to a context that also times the sieve and the marking apart
(GB_CFG_KERNEL_TIMES, which keeps one stream), and equal there to the sum of that split.  (The
reference reports only wall time, goldbach.cu:702-758.)"""
import time

import pytest

pytestmark = pytest.mark.gpu

LIMIT = 10**9
SEG = 1 << 24
STAT_KEYS = ("evens", "phase2", "failures", "first_fail_n")


def test_plain_kernel_time_comes_from_the_device_clock(gb):
    evens = (LIMIT - 4) // 2 + 1
    segments = -(-evens // SEG)
    c = gb.Context(0, max_seg_evens=SEG)
    try:
        c.verify_range(4, LIMIT, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=SEG)   # warm
        t0 = time.perf_counter()
        st = c.verify_range(4, LIMIT, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=SEG)
        wall_ms = (time.perf_counter() - t0) * 1e3
    finally:
        c.close()
    assert st["evens"] == evens and st["failures"] == 0 and st["phase2"] == 0
    assert 0 < st["kernel_ms"] <= wall_ms

    t = gb.Context(0, max_seg_evens=SEG, kernel_times=True)
    try:
        t.kernel_times(reset=True)
        st2 = t.verify_range(4, LIMIT, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=SEG)
        kt = t.kernel_times(reset=True)
        again = t.kernel_times(reset=False)
    finally:
        t.close()
    assert {k: st2[k] for k in STAT_KEYS} == {k: st[k] for k in STAT_KEYS}
    assert kt["segments"] == segments and kt["evens"] == evens
    assert kt["sieve_ms"] > 0 and kt["mark_ms"] > 0
    total = kt["sieve_ms"] + kt["mark_ms"] + kt["other_ms"]
    assert total == pytest.approx(st2["kernel_ms"], rel=1e-9, abs=1e-9)
    # the split's events and the device clock agree: what is left is small and not negative
    assert -0.05 * total < kt["other_ms"] < total
    assert again["segments"] == 0   # reset=True cleared the counts
