"""The batched form's mod-6 wheel (q held as 6k+1 / 6k+5 streams, even n in three
classes n mod 6) against the oracle: every range start and end residue mod 6,
ranges of one or two n, tile edges, the smallest tiles, the first tiles (where
stream bits stand for q <= 1 and 6 = 3 + 3 lies outside the wheel) and the 2^44
edge above which windows go back to the odd-only Miller-Rabin tiles.  With
p_min statistics the table-driven class marking reads every stream bit a
partition scan touches, so sum/max/argmax of p_min pin the wheel's sieve; the
unrolled path is checked against it.  Reference semantics: goldbach.cu:288-311
(Phase 1), :362-373 (Phase 2), cpu_goldbach.cpp:19-22 (p ascending)."""
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n")


def check_range(ctx, gb, orc, start, limit, p_small=10**6, tile_log2=0):
    ref = orc.verify_range(start, limit, 1 << 22, p_small)
    got = ctx.verify_range(start, limit, p_small, form=gb.GB_FORM_BATCHED, tile_log2=tile_log2, want_pmin=True)
    for k in KEYS:
        assert got[k] == ref[k], (start, limit, p_small, tile_log2, k, got[k], ref[k])
    fast = ctx.verify_range(start, limit, p_small, form=gb.GB_FORM_BATCHED, tile_log2=tile_log2)
    for k in ("evens", "phase2", "failures", "first_fail_n"):
        assert fast[k] == got[k], (start, limit, k)


@pytest.mark.parametrize("s_mod", [0, 2, 4])
@pytest.mark.parametrize("e_mod", [0, 2, 4])
def test_start_and_end_residues(ctx, gb, orc, s_mod, e_mod):
    start = 6 * 166_667 + s_mod
    limit = 6 * 500_000 + 6 * 1003 + e_mod
    check_range(ctx, gb, orc, start, limit, tile_log2=12)


@pytest.mark.parametrize("start,limit", [(4, 4), (4, 6), (6, 6), (4, 8), (8, 8), (10, 12), (4, 100),
                                         (4, 30_000), (8_440, 8_460), (12_285, 12_400)])
def test_first_tiles_and_tiny_ranges(ctx, gb, orc, start, limit):
    check_range(ctx, gb, orc, start, limit, tile_log2=12)
    check_range(ctx, gb, orc, start, limit)


@pytest.mark.parametrize("tile_log2", [12, 13, 16, 19, 20])
def test_tile_edges(ctx, gb, orc, tile_log2):
    cls = 9 << (tile_log2 - 4)            # n per class per tile: a tile spans 6 * cls
    for k in (1, 2, 3):
        check_range(ctx, gb, orc, 4, 6 * cls * k - 2, tile_log2=tile_log2)
        check_range(ctx, gb, orc, 4, 6 * cls * k, tile_log2=tile_log2)


def test_small_p_small_through_the_wheel(ctx, gb, orc):
    for p_small in (0, 2, 3, 5, 7, 31, 8191, 8193):
        check_range(ctx, gb, orc, 4, 3 * 10**6, p_small=p_small, tile_log2=14)


def test_failure_hook_through_the_wheel(gb, orc):
    # Phase 2 cut at 7: the first n without a partition p <= 7 is 98 (goldbach.cu:476-480)
    with gb.Context(0, phase2_limit=7) as c:
        st = c.verify_range(4, 10**5, 3, form=gb.GB_FORM_BATCHED, want_pmin=True)
        ref = orc.verify_range(4, 10**5, 1 << 22, 3, phase2_limit=7)
        for k in KEYS + ("first_fail_n",):
            assert st[k] == ref[k], k
        assert st["first_fail_n"] == 98


def test_windows_around_the_wheel_edge(ctx, orc):
    edge = 2**44 - 2**24
    below = [(edge - 2 * 5000, 4000), (10**13 + 2, 3000), (4, 2000)]
    above = below + [(edge + 2, 3000)]                  # one window above: all take the odd-only tiles
    for wins in (below, above):
        per, tot = ctx.verify_windows(wins, 10**6, want_pmin=True)
        for (a, e), got in zip(wins, per):
            st, _, _ = orc.verify_segment(a, e, 10**6, want_bits=False, want_pmin=False)
            for k in KEYS:
                assert got[k] == st[k], (a, e, k)


@pytest.mark.parametrize("tile_log2", [12, 16, 19])
def test_two_workgroups_per_cu(ctx, gb, orc, tile_log2):
    # half-size workgroups, two per CU (the default below 2^44 is tile_log2 19 this way)
    cls = 9 << (tile_log2 - 4)
    for start, limit in ((4, 6 * cls * 3), (1_000_002, 1_000_002 + 6 * cls * 2 + 10)):
        ref = orc.verify_range(start, limit, 1 << 22, 10**6)
        got = ctx.verify_range(start, limit, 10**6, form=gb.GB_FORM_BATCHED, tile_log2=tile_log2, wg_per_cu=2,
                               want_pmin=True)
        for k in KEYS:
            assert got[k] == ref[k], (tile_log2, start, k)
        fast = ctx.verify_range(start, limit, 10**6, form=gb.GB_FORM_BATCHED, tile_log2=tile_log2, wg_per_cu=2)
        assert (fast["evens"], fast["phase2"], fast["failures"]) == (ref["evens"], 0, 0)
    with pytest.raises(gb.GoldbachError):
        ctx.verify_range(4, 10**6, 10**6, form=gb.GB_FORM_BATCHED, tile_log2=20, wg_per_cu=2)
