"""Device verifier parity against the oracle: per-segment verified bitmaps,
per-n minimal partitions, Phase-2 and failure counts (reference semantics:
src/goldbach.cu:288-311 Phase 1, :362-373 Phase 2, :420-485 segment loop),
the batched persistent form against the plain form, windows that need
Miller-Rabin, and single-number minimal partitions (single_check's job)."""
import numpy as np
import pytest

from _helpers import unpack

pytestmark = pytest.mark.gpu

SEGMENTS = [
    (4, 1, 10**6), (4, 2, 10**6), (6, 1, 10**6), (4, 63, 10**6), (4, 64, 10**6), (4, 65, 10**6),
    (4, 1000, 0), (4, 1000, 2), (4, 1000, 3), (4, 1000, 5), (4, 1000, 7), (4, 5000, 100),
    (4, 200_000, 10**6), (8190, 5000, 10**6), (8194, 5000, 10**6), (10**6, 100_000, 1000),
    (10**9, 1 << 17, 10**6), (10**9 + 2, 12_345, 8193), (10**9 + 2, 12_345, 8194),
    (10**12 - 2 * (1 << 20), 1 << 20, 10**6), (10**12, 3000, 4 * 10**9),
    (10**15, 20_000, 10**6), (10**18 + 2, 10_000, 10**6),
    (2**64 - 2 - 2 * 4095, 4096, 10**6),
]


@pytest.mark.parametrize("A,count,p_small", SEGMENTS)
def test_segment_matches_oracle(ctx, orc, A, count, p_small):
    r = ctx.verify_segment(A, count, p_small, want_pmin_array=True)
    st, bits, pmin = orc.verify_segment(A, count, p_small)
    assert np.array_equal(r.verified, bits)
    assert np.array_equal(r.pmin, pmin)
    for k in ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin", "argmax_n"):
        assert r.stats[k] == st[k], k
    # without per-n outputs the unrolled marking path runs (p_small >= 8193)
    f = ctx.verify_segment(A, count, p_small)
    assert np.array_equal(f.verified, bits)
    for k in ("evens", "phase2", "failures", "first_fail_n"):
        assert f.stats[k] == st[k], k


def test_segment_slow_path_record(ctx, orc):
    # p_min(3325581707333960528) = 9781 lies beyond the 8193 the window serves
    n = 3325581707333960528
    A = n - 2 * 300
    r = ctx.verify_segment(A, 700, 10**6, want_pmin_array=True)
    st, bits, pmin = orc.verify_segment(A, 700, 10**6)
    assert int(r.pmin[300]) == 9781
    assert np.array_equal(r.pmin, pmin) and np.array_equal(r.verified, bits)
    assert (r.stats["max_pmin"], r.stats["argmax_n"]) == (9781, n)
    # with p_small below the record the slow path reports Phase 2 for it
    r2 = ctx.verify_segment(A, 700, 9000, want_pmin_array=True)
    st2, bits2, _ = orc.verify_segment(A, 700, 9000)
    assert np.array_equal(r2.verified, bits2) and r2.stats["phase2"] == st2["phase2"] == 1


def test_segment_bad_arguments(ctx, gb):
    for args in [(5, 10, 100), (2, 10, 100), (4, 0, 100), (2**64 - 2, 2, 100)]:
        with pytest.raises(gb.GoldbachError) as e:
            ctx.verify_segment(*args)
        assert e.value.status == gb.GB_EINVAL
    with pytest.raises(gb.GoldbachError) as e:
        ctx.verify_segment(4, (1 << 27) + 1, 100)
    assert e.value.status == gb.GB_ERANGE


def test_failure_hook(gb, orc):
    # Phase 2 cut to p <= 7: the first failing n is the record n = 98
    with gb.Context(0, phase2_limit=7) as c:
        r = c.verify_segment(4, 499, 3, want_pmin_array=True)
        st, bits, pmin = orc.verify_segment(4, 499, 3, phase2_limit=7)
        assert r.stats["first_fail_n"] == 98
        for k in ("phase2", "failures", "first_fail_n", "sum_pmin"):
            assert r.stats[k] == st[k]
        assert np.array_equal(r.pmin, pmin) and np.array_equal(r.verified, bits)
        for form in (gb.GB_FORM_PLAIN, gb.GB_FORM_BATCHED):
            rr = c.verify_range(4, 1000, 3, form=form)
            assert rr["first_fail_n"] == 98 and rr["failures"] > 0


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_range_1e8_known_sums(ctx, gb, form):
    f = gb.GB_FORM_PLAIN if form == "plain" else gb.GB_FORM_BATCHED
    st = ctx.verify_range(4, 10**8, 10**6, form=f, want_pmin=True)
    assert st["evens"] == 49_999_999 and st["phase2"] == 0 and st["failures"] == 0
    assert st["sum_pmin"] == 1_511_603_116
    assert (st["max_pmin"], st["argmax_n"]) == (1093, 60119912)


@pytest.mark.parametrize("tile_log2", [12, 14, 16, 18, 19, 20])
def test_batched_tile_sizes_agree(ctx, gb, tile_log2):
    st = ctx.verify_range(10**6, 10**8 + 10**6, 1000, form=gb.GB_FORM_BATCHED, tile_log2=tile_log2,
                          want_pmin=True)
    ref = ctx.verify_range(10**6, 10**8 + 10**6, 1000, form=gb.GB_FORM_PLAIN, seg_evens=10**7, want_pmin=True)
    for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
        assert st[k] == ref[k], k


def test_range_matches_oracle_with_start(ctx, gb, orc):
    # validation_gpu.sh 1.1/1.3 settings: --seg-size=1e6 --p-small=1e5, start 1e6
    ref = orc.verify_range(1_000_000, 10**9, 10**6, 100_000)
    for form in (gb.GB_FORM_PLAIN, gb.GB_FORM_BATCHED):
        st = ctx.verify_range(1_000_000, 10**9, 100_000, form=form, seg_evens=10**6, want_pmin=True)
        for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
            assert st[k] == ref[k], (form, k)


def test_range_small_p_small_counts_phase2(ctx, gb, orc):
    ref = orc.verify_range(4, 2 * 10**7, 1 << 22, 31)
    for form in (gb.GB_FORM_PLAIN, gb.GB_FORM_BATCHED):
        st = ctx.verify_range(4, 2 * 10**7, 31, form=form, seg_evens=1 << 22, want_pmin=True)
        for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
            assert st[k] == ref[k], (form, k)
    assert ref["phase2"] > 0


def test_full_size_1e11_plain_equals_batched(ctx, gb):
    # the plain per-segment form at 2^27 evens per segment and the batched form
    # share no kernel launch geometry; their per-range sums must agree exactly
    b = ctx.verify_range(4, 10**11, 10**6, form=gb.GB_FORM_BATCHED, want_pmin=True)
    p = ctx.verify_range(4, 10**11, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=1 << 27, want_pmin=True)
    assert b["evens"] == 49_999_999_999 and b["failures"] == 0 and b["phase2"] == 0
    for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
        assert b[k] == p[k], k
    # the same range through the unrolled marking path (no p_min statistics)
    for form in (gb.GB_FORM_BATCHED, gb.GB_FORM_PLAIN):
        f = ctx.verify_range(4, 10**11, 10**6, form=form, seg_evens=1 << 27)
        assert (f["evens"], f["phase2"], f["failures"]) == (49_999_999_999, 0, 0)
    # a small P_SMALL forces Phase 2 counts that both paths must agree on
    q = ctx.verify_range(4, 10**9, 100, form=gb.GB_FORM_BATCHED)
    r = ctx.verify_range(4, 10**9, 100, form=gb.GB_FORM_BATCHED, want_pmin=True)
    assert q["phase2"] == r["phase2"] > 0


def test_windows_match_oracle(ctx, orc):
    wins = [(10**15, 20_000), (10**18 + 2, 5_000), (1234567890123456, 7_777), (2**64 - 2 - 2 * 3999, 4000),
            (4, 3000), (3325581707333960528 - 600, 700),
            (2**44 - 20_000, 20_000),                      # q and n on both sides of 2^44
            (2**44 - 2**21, (1 << 20) + 5)]                # tiles below 2^44 (wheel) and above (Baillie-PSW)
    per, tot = ctx.verify_windows(wins, 10**6, want_pmin=True)
    evens = 0
    for (a, e), got in zip(wins, per):
        st, _, _ = orc.verify_segment(a, e, 10**6, want_bits=False, want_pmin=False)
        for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
            assert got[k] == st[k], (a, k)
        evens += e
    assert tot["evens"] == evens


def test_many_ragged_windows_span_several_launches(ctx, orc):
    # more windows than one launch takes pieces (64), of ragged lengths down to a single even n,
    # below the wheel bound, across 2^44 - 2^24 and 2^44, and in the Baillie-PSW range; one window
    # listed twice and two overlapping
    rng = np.random.default_rng(7)
    wins = []
    for lo, hi, k, emax in ((4, 2**40, 60, 3000), (2**44 - 2**25, 2**44 + 2**22, 30, 3000),
                            (10**15, 2**64 - 2 - 2 * 3000, 60, 2000)):
        a = rng.integers(lo // 2, hi // 2, size=k, dtype=np.uint64) * np.uint64(2)
        e = rng.integers(1, emax + 1, size=k)
        wins += [(int(x), int(y)) for x, y in zip(a, e)]
    wins += [wins[5], (wins[70][0] + 2 * 100, 500), (4, 1), (6, 1)]
    per, tot = ctx.verify_windows(wins, 10**6, want_pmin=True)
    for (a, e), got in zip(wins, per):
        st, _, _ = orc.verify_segment(a, e, 10**6, want_bits=False, want_pmin=False)
        for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
            assert got[k] == st[k], (a, e, k)
    assert tot["evens"] == sum(e for _, e in wins)
    # without p_min statistics the fast marking path runs; its counts must not change
    per2, _ = ctx.verify_windows(wins, 10**6)
    for got, want in zip(per2, per):
        assert (got["evens"], got["phase2"], got["failures"]) == (want["evens"], want["phase2"], want["failures"])


@pytest.mark.parametrize("stride", [1, 2, 3, 5, 6, 8, 64])
def test_mr_proof_stride_leaves_the_counts_alone(ctx, gb, orc, stride):
    # odd-only tiles above 2^44 prove the survivors of one window word in `stride`; an n that no
    # proven q resolves gets a second pass over the pre-sieve survivors of the unproven words, and
    # what that leaves goes to the slow path, which decides it exactly from p = 3 (at stride 64
    # about a third of the n are open: more than a tile's open list holds).  P_SMALL = 8193
    # makes the record n 3325581707333960528 (p_min 9781) a Phase-2 case.
    rec = 3325581707333960528
    for a, e, ps in ((10**15, 30_000, 10**6), (2**63 + 2, 20_000, 10**6), (2**64 - 2 - 2 * 19_999, 20_000, 10**6),
                     (rec - 2 * 5000, 10_000, 8193)):
        st, _, _ = orc.verify_segment(a, e, ps, want_bits=False, want_pmin=False)
        got = ctx.verify_range(a, a + 2 * (e - 1), ps, mr_proof_stride=stride)
        assert (got["evens"], got["phase2"], got["failures"]) == (st["evens"], st["phase2"], st["failures"]), (a, stride)
        if a == rec - 2 * 5000:
            assert st["phase2"] >= 1
    with pytest.raises(gb.GoldbachError):
        ctx.verify_range(10**15, 10**15 + 1000, 10**6, mr_proof_stride=65)


def test_autotune_picks_a_valid_tuning(ctx, gb):
    want = ctx.verify_range(4, 10**11, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=1 << 27)
    t, rate = ctx.autotune(4, 10**11, 10**6, form=gb.GB_FORM_BATCHED)
    assert 16 <= t["tile_log2"] <= 20 and t["wg_per_cu"] in (1, 2) and rate > 1e10
    assert t["lane_item_div"] in (0, 64, 128, 256) and t["wave_item_div"] in (0, 1024, 2048)
    assert ctx.autotune(4, 10**11, 10**6, form=gb.GB_FORM_BATCHED)[0] == t   # cached
    got = ctx.verify_range(4, 10**11, 10**6, form=gb.GB_FORM_BATCHED, autotune=True)
    assert (got["evens"], got["phase2"], got["failures"]) == (want["evens"], want["phase2"], want["failures"])
    tp, rate_p = ctx.autotune(4, 10**11, 10**6, form=gb.GB_FORM_PLAIN)
    assert tp["seg_evens"] in [1 << k for k in range(24, 28)] and rate_p > 1e9
    got = ctx.verify_range(4, 10**11, 10**6, form=gb.GB_FORM_PLAIN, autotune=True)
    assert (got["evens"], got["phase2"], got["failures"]) == (want["evens"], want["phase2"], want["failures"])
    # a range too short to sample keeps the defaults
    short, r0 = ctx.autotune(4, 10**6, 10**6)
    assert (short["tile_log2"], short["wg_per_cu"], r0) == (19, 2, 0)


def high_windows(count, lo_evens, hi_evens, seed=0):
    """Windows of uneven length starting at even offsets log-uniform in [1e15, 1.8e19]."""
    rng = np.random.default_rng(seed)
    starts = np.exp(rng.uniform(np.log(1e15), np.log(1.8e19), size=count))
    lens = rng.integers(lo_evens, hi_evens + 1, size=count)
    return [(int(a / 2) * 2, int(e)) for a, e in zip(starts, lens)]


def test_high_windows_draw_matches_oracle(ctx, orc):
    # the high-window workload's draw (seed 0), lengths scaled down so the oracle finishes in seconds
    wins = high_windows(64, 1000, 30_000)
    per, tot = ctx.verify_windows(wins, 10**6, want_pmin=True)
    for (a, e), got in zip(wins, per):
        st, _, _ = orc.verify_segment(a, e, 10**6, want_bits=False, want_pmin=False)
        for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
            assert got[k] == st[k], (a, e, k)
    assert tot["evens"] == sum(e for _, e in wins) and tot["failures"] == 0
    # the same windows one by one through the plain per-segment path
    for (a, e), got in zip(wins[:8], per[:8]):
        r = ctx.verify_segment(a, e, 10**6, want_pmin=True)
        for k in ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n"):
            assert r.stats[k] == got[k], (a, k)


def test_min_partition_matches_oracle(ctx, orc):
    rng = np.random.default_rng(1)
    n = (np.exp(rng.uniform(np.log(1e6), np.log(2.0**64 - 2), size=3000)) / 2).astype(np.uint64) * np.uint64(2)
    n = np.concatenate([n, np.array([4, 6, 8, 10, 98, 2**64 - 2, 3325581707333960528, 10**19], dtype=np.uint64)])
    got = ctx.min_partition(n)
    want = orc.min_partition_batch(n)
    assert np.array_equal(got, want)
    # q = n - 3 a strong pseudoprime to the first 8 and to the first 11 bases
    for psi in (341550071728321, 3825123056546413051, 2152302898747, 3474749660383):
        m = np.array([psi + 3, psi + 5, psi + 7], dtype=np.uint64)
        assert np.array_equal(ctx.min_partition(m), orc.min_partition_batch(m))
    # odd or too small: no partition reported
    assert list(ctx.min_partition([2, 3, 7, 0])) == [0, 0, 0, 0]
    # each result is the same alone and in the mix
    for i in (0, 17, 2999):
        assert ctx.min_partition(n[i:i + 1])[0] == got[i]


def _range_fixtures():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "range_sums.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_range_sums_equal_fixture(ctx, gb, form):
    # [4, 1e9], [4, 1e10], [1e10, 1e10 + 2^31], [1e12 - 2^31, 1e12] and the whole headline range
    # [4, 1e12] against the oracle's statistics (sum, maximum and argmax of p_min, Phase-2 and
    # failure counts); the maxima from 4 are published records (721013438: 1789,
    # 7473202036: 1877, 599533546358: 3457)
    f = gb.GB_FORM_BATCHED if form == "batched" else gb.GB_FORM_PLAIN
    for r in _range_fixtures():
        st = ctx.verify_range(r["start"], r["limit"], r["p_small"], form=f, want_pmin=True)
        for k in ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin", "argmax_n"):
            assert st[k] == r[k], (r["start"], r["limit"], k)


@pytest.mark.parametrize("lane_div,wave_div", [(24, 1024), (48, 2048), (16, 4096)])
def test_work_item_bounds_leave_the_answers_alone(ctx, gb, lane_div, wave_div):
    # the medium primes' work-item bounds (which primes are cut into lane items, which into wave
    # items, which get a thread each) move work between the sieve's code paths; the statistics of
    # [4, 1e10] and of the window near 1e12 must not move
    for r in _range_fixtures():
        if r["limit"] - r["start"] > 2 * 10**10:
            continue
        st = ctx.verify_range(r["start"], r["limit"], r["p_small"], form=gb.GB_FORM_BATCHED, want_pmin=True,
                              lane_item_div=lane_div, wave_item_div=wave_div)
        for k in ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin", "argmax_n"):
            assert st[k] == r[k], (r["start"], r["limit"], lane_div, wave_div, k)


def test_work_item_bounds_are_checked(ctx, gb):
    for kw in ({"lane_item_div": 4}, {"lane_item_div": 512}, {"wave_item_div": 100}):
        with pytest.raises(gb.GoldbachError):
            ctx.verify_range(4, 10**8, 10**6, form=gb.GB_FORM_BATCHED, **kw)
