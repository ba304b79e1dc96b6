"""The batched form over ranges of any length: launches of bounded size on the tile grid of
their kernel (the reference keeps fixed per-segment buffers whatever the range,
src/goldbach.cu:378-417, README.md:74-75), progress after every launch
(goldbach.cu:702-758), the stop at the first failure (goldbach.cu:420,476-480), the split
where the mod-6 wheel stops serving, and the device-memory budget every allocation counts
against."""
import ctypes
import threading
import time

import pytest

pytestmark = pytest.mark.gpu

WHEEL_N_MAX = (1 << 44) - (1 << 24)
CHUNK_TILES = 1 << 17
WHEEL_SPAN = 6 * (9 << 15)   # n spanned by a default wheel tile: 3 classes of 9 * 2^15 even n
KEYS = ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin", "argmax_n")


def covers(plan, start, limit):
    """the chunks tile [start, limit] in order, without gaps or overlaps"""
    assert plan[0]["first_n"] == start and plan[-1]["last_n"] == limit
    for a, b in zip(plan, plan[1:]):
        assert b["first_n"] == a["last_n"] + 2
    return True


def test_plan_of_1e13_is_bounded_wheel_launches(ctx, gb):
    plan = ctx.range_plan(4, 10**13)
    assert covers(plan, 4, 10**13)
    assert all(c["kernel"] == gb.GB_KERNEL_WHEEL and c["tiles"] <= CHUNK_TILES for c in plan)
    assert (plan[0]["tile_log2"], plan[0]["wg_per_cu"]) == (19, 2)
    # whole chunks are whole numbers of tiles: every chunk after the first starts a tile
    assert all(c["first_n"] % WHEEL_SPAN == 0 for c in plan[1:])
    assert len(plan) == -(-(10**13 // WHEEL_SPAN + 1) // CHUNK_TILES)


def test_plan_splits_at_the_wheel_bound(ctx, gb):
    plan = ctx.range_plan(4, 2 * 10**13)
    assert covers(plan, 4, 2 * 10**13)
    wheel = [c for c in plan if c["kernel"] == gb.GB_KERNEL_WHEEL]
    fused = [c for c in plan if c["kernel"] == gb.GB_KERNEL_FUSED]
    assert wheel and fused and plan == wheel + fused
    assert wheel[-1]["last_n"] == WHEEL_N_MAX and fused[0]["first_n"] == WHEEL_N_MAX + 2
    assert (fused[0]["tile_log2"], fused[0]["wg_per_cu"]) == (20, 1)
    assert all(c["tiles"] <= CHUNK_TILES for c in plan)
    # the plain form cuts segments of seg_evens
    p = ctx.range_plan(4, 10**9, form=gb.GB_FORM_PLAIN, seg_evens=1 << 27)
    assert covers(p, 4, 10**9) and len(p) == 4 and p[0]["last_n"] == 4 + 2 * ((1 << 27) - 1)


def test_range_across_the_wheel_bound_matches_oracle(ctx, gb, orc):
    a, b = WHEEL_N_MAX - (1 << 26), WHEEL_N_MAX + (1 << 26)
    plan = ctx.range_plan(a, b)
    assert [c["kernel"] for c in plan] == [gb.GB_KERNEL_WHEEL, gb.GB_KERNEL_FUSED]
    st = ctx.verify_range(a, b, 10**6, want_pmin=True)
    ref = orc.verify_range(a, b, 1 << 27, 10**6)
    for k in KEYS:
        assert st[k] == ref[k], k


def test_1e13_within_a_budget_per_tile_records_would_exceed(gb):
    # one 64-byte record per wheel tile over [4, 1e13] would take 362 MB; the context may hold 256 MiB
    tiles = 10**13 // WHEEL_SPAN + 1
    assert tiles * 64 > 256 << 20
    with gb.Context(0, mem_budget=256 << 20) as c:
        b = c.verify_range(4, 10**13, 10**6)
        p = c.verify_range(4, 10**13, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=1 << 27)
    assert b["evens"] == 5 * 10**12 - 1 and b["failures"] == 0 and b["phase2"] == 0
    for k in ("evens", "phase2", "failures", "first_fail_n"):
        assert b[k] == p[k], k


def test_budget_refuses_what_it_cannot_hold(gb):
    with gb.Context(0, mem_budget=256 << 20) as c:
        with pytest.raises(gb.GoldbachError) as e:
            c.count_partitions(1 << 35, 4)          # two 2 GiB bitsets
        assert e.value.status == gb.GB_ENOMEM and "budget" in e.value.message
        # the context still works after the refusal
        assert c.verify_range(4, 10**6, 10**6)["evens"] == (10**6 - 4) // 2 + 1


def test_progress_moves_during_a_batched_run(ctx):
    total = (10**13 - 4) // 2 + 1
    prog = ctypes.c_uint64(0)
    seen = set()
    out = {}
    th = threading.Thread(target=lambda: out.update(ctx.verify_range(4, 10**13, 10**6, progress=prog)))
    th.start()
    while th.is_alive():
        seen.add(prog.value)
        time.sleep(0.001)
    th.join()
    assert out["evens"] == total and prog.value == total
    assert any(0 < v < total for v in seen), sorted(seen)[:5]


def test_batched_range_stops_at_the_first_failing_launch(gb):
    # Phase 2 cut at 7, P_SMALL 3: n = 98 fails, in the first launch; nothing after it runs
    with gb.Context(0, phase2_limit=7) as c:
        plan = c.range_plan(4, 5 * 10**11, 3)
        assert len(plan) >= 3
        st = c.verify_range(4, 5 * 10**11, 3)
    first = plan[0]
    assert st["first_fail_n"] == 98
    assert st["evens"] == (first["last_n"] - first["first_n"]) // 2 + 1


@pytest.mark.parametrize("tile_log2", [0, 12])
def test_multi_launch_range_near_1e12_equals_fixture(ctx, gb, tile_log2):
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "range_sums.json")) as f:
        fx = {(r["start"], r["limit"]): r for r in json.load(f)}
    a, b = 10**12 - 2**31, 10**12
    if (a, b) not in fx:
        pytest.skip("range_sums.json has no [1e12 - 2^31, 1e12] entry")
    plan = ctx.range_plan(a, b, tile_log2=tile_log2)
    assert covers(plan, a, b) and (tile_log2 == 0 or len(plan) >= 2)
    st = ctx.verify_range(a, b, 10**6, tile_log2=tile_log2, want_pmin=True)
    for k in KEYS:
        assert st[k] == fx[(a, b)][k], k


def test_min_partition_and_serving_under_a_tiny_phase2_bound(gb, orc):
    # phase2_limit 7 leaves q = n - p as small as 3 to the exact 64-bit test
    ns = list(range(4, 1001, 2))
    want = [orc.min_partition(n, 7) for n in ns]
    with gb.Context(0, phase2_limit=7) as c:
        assert [int(x) for x in c.min_partition(ns)] == want
        with c.server() as s:
            assert [int(x) for x in s.run(ns)] == want
        # below 2^64 big_check's search reaches p_max, not the context's Phase-2 bound
        p, _ = c.big_min_partition(98)
        assert p == orc.min_partition(98, 10**7) == 19


def test_plain_failure_chunk_is_logged_and_resumed(gb, tmp_path):
    # plain segments far shorter than the log's chunks: the failing chunk stops after its
    # failing segment, is recorded short, and a second call resumes past it
    path = str(tmp_path / "p.log")
    with gb.Context(0, phase2_limit=7) as c:
        st, info = c.verify_range_logged(4, 10**5, 3, path, chunk_evens=1000, form=gb.GB_FORM_PLAIN, seg_evens=16)
        assert st["first_fail_n"] == 98 and info["complete"] == 1 and info["done_chunks"] == 1
        st2, info2 = c.verify_range_logged(4, 10**5, 3, path, chunk_evens=1000, form=gb.GB_FORM_PLAIN,
                                           seg_evens=16)
        assert info2["resumed_chunks"] == 1 and st2["first_fail_n"] == 98 and st2["evens"] == st["evens"]
    _, recs, torn = gb.read_log(path)
    assert len(recs) == 1 and not torn and recs[0]["evens"] < 1000
