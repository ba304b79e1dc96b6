"""Resumable runs on the device (gb_verify_range_logged): a run cut into logged
chunks gives the statistics of the uncut run; a run stopped after some chunks,
or whose log lost its tail in a crash, resumes where its intact records end
and reaches the same totals; a log written for another run is refused; a
chunk with a failure ends the run as the reference stops at its first failure
(goldbach.cu:476-480); the CLI's --log resumes instead of --start
(goldbach.cu:581,604-609,689)."""
import fcntl
import os
import subprocess

import pytest

from _helpers import PKG_DIR

pytestmark = pytest.mark.gpu
GOLDBACH = os.path.join(PKG_DIR, "bin", "goldbach")
KEYS = ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin", "argmax_n")


def same(a, b):
    for k in KEYS:
        assert a[k] == b[k], (k, a[k], b[k])


def test_logged_equals_unlogged(ctx, gb, tmp_path):
    ref = ctx.verify_range(4, 10**8, 10**6, want_pmin=True)
    assert ref["sum_pmin"] == 1_511_603_116                     # pinned by the oracle to 1e8
    st, info = ctx.verify_range_logged(4, 10**8, 10**6, str(tmp_path / "a.log"), chunk_evens=7_000_000,
                                       want_pmin=True)
    same(st, ref)
    assert info["complete"] == 1 and info["resumed_chunks"] == 0 and info["done_chunks"] == info["total_chunks"] == 8
    h, recs, torn = gb.read_log(str(tmp_path / "a.log"))
    assert not torn and len(recs) == 8 and h["want_pmin"] == 1
    assert sum(r["evens"] for r in recs) == ref["evens"]
    # plain form and other tiles give the same records
    st2, _ = ctx.verify_range_logged(4, 10**8, 10**6, str(tmp_path / "b.log"), chunk_evens=7_000_000,
                                     form=gb.GB_FORM_PLAIN, seg_evens=1 << 22, want_pmin=True)
    same(st2, ref)
    assert open(tmp_path / "a.log").read() == open(tmp_path / "b.log").read()


def test_stop_and_resume(ctx, tmp_path):
    path = str(tmp_path / "r.log")
    ref = ctx.verify_range(10**9, 3 * 10**9, 10**6)
    st, info = ctx.verify_range_logged(10**9, 3 * 10**9, 10**6, path, chunk_evens=1 << 27, max_chunks=3)
    assert info["complete"] == 0 and info["done_chunks"] == 3 and st["evens"] == 3 << 27
    st, info = ctx.verify_range_logged(10**9, 3 * 10**9, 10**6, path, chunk_evens=1 << 27, form=0)
    assert info["resumed_chunks"] == 3 and info["resumed_evens"] == 3 << 27 and info["complete"] == 1
    same(st, ref)
    # a finished log answers without new chunks
    st, info = ctx.verify_range_logged(10**9, 3 * 10**9, 10**6, path, chunk_evens=1 << 27)
    assert info["resumed_chunks"] == info["total_chunks"] and st["kernel_ms"] == 0
    same(st, ref)


def test_damaged_tail_is_recomputed(ctx, gb, tmp_path):
    path = str(tmp_path / "t.log")
    ref = ctx.verify_range(4, 2 * 10**7, 10**5, want_pmin=True)
    ctx.verify_range_logged(4, 2 * 10**7, 10**5, path, chunk_evens=10**6, want_pmin=True)
    lines = open(path).read().splitlines(keepends=True)
    # a crash mid-write: record 6 half written; record 4 damaged on disk
    open(path, "w").write("".join(lines[:7]) + lines[7][:25])
    st, info = ctx.verify_range_logged(4, 2 * 10**7, 10**5, path, chunk_evens=10**6, want_pmin=True)
    assert info["truncated"] == 1 and info["resumed_chunks"] == 6
    same(st, ref)
    bad = lines[5].replace("chunk=4 ", "chunk=4  ")
    open(path, "w").write("".join(lines[:5]) + bad + "".join(lines[6:]))
    st, info = ctx.verify_range_logged(4, 2 * 10**7, 10**5, path, chunk_evens=10**6, want_pmin=True)
    assert info["truncated"] == 1 and info["resumed_chunks"] == 4
    same(st, ref)
    _, recs, torn = gb.read_log(path)
    assert not torn and len(recs) == info["total_chunks"]
    assert open(path).read() == "".join(lines)


def test_log_of_another_run_is_refused(ctx, gb, tmp_path):
    path = str(tmp_path / "o.log")
    ctx.verify_range_logged(4, 10**6, 10**5, path, chunk_evens=10**5, max_chunks=1)
    before = open(path).read()
    for kw in ({"limit": 2 * 10**6}, {"p_small": 10**4}, {"chunk_evens": 2 * 10**5}, {"want_pmin": True},
               {"start": 6}):
        args = {"start": 4, "limit": 10**6, "p_small": 10**5, "chunk_evens": 10**5}
        args.update(kw)
        with pytest.raises(gb.GoldbachError) as e:
            ctx.verify_range_logged(args["start"], args["limit"], args["p_small"], path,
                                    chunk_evens=args["chunk_evens"], want_pmin=args.get("want_pmin", False))
        assert e.value.status == gb.GB_EINVAL and "another run" in e.value.message, kw
    assert open(path).read() == before
    open(path, "w").write("not a log\n")
    with pytest.raises(gb.GoldbachError):
        ctx.verify_range_logged(4, 10**6, 10**5, path, chunk_evens=10**5)
    # the range adjustments of the reference apply before the comparison: 3 -> 4, 10^6 + 1 -> 10^6
    path2 = str(tmp_path / "n.log")
    ctx.verify_range_logged(3, 10**6 + 1, 10**5, path2, chunk_evens=10**5, max_chunks=1)
    st, info = ctx.verify_range_logged(4, 10**6, 10**5, path2, chunk_evens=10**5)
    assert info["resumed_chunks"] == 1 and st["evens"] == (10**6 - 4) // 2 + 1


def test_failure_ends_the_run(gb, tmp_path):
    # Phase 2 cut at 7 with P_SMALL 3: the first n without a partition p <= 7 is 98
    path = str(tmp_path / "f.log")
    with gb.Context(0, phase2_limit=7) as c:
        st, info = c.verify_range_logged(4, 10**5, 3, path, chunk_evens=20)
        assert st["first_fail_n"] == 98 and st["failures"] >= 1
        assert info["complete"] == 1 and info["done_chunks"] == 3     # 98 lies in chunk 2: [84, 122]
        st2, info2 = c.verify_range_logged(4, 10**5, 3, path, chunk_evens=20)
        assert info2["resumed_chunks"] == 3 and st2["first_fail_n"] == 98 and st2["kernel_ms"] == 0
    _, recs, _ = gb.read_log(path)
    assert len(recs) == 3 and recs[2]["first_fail_n"] == 98 and recs[2]["first_n"] == 84


def test_cli_log_flag(tmp_path):
    path = str(tmp_path / "cli.log")
    args = [GOLDBACH, "4000000000", "--seg-size=10000000", f"--log={path}", "--log-chunk=300000000"]
    r1 = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r1.returncode == 0 and "satisfy Goldbach. ✓" in r1.stdout, r1.stdout + r1.stderr
    assert "Run log        : " + path + " (0 of 7 chunks resumed, 0 evens)" in r1.stdout
    lines = open(path).read().splitlines(keepends=True)
    open(path, "w").write("".join(lines[:4]))                      # header + 3 chunks survive
    r2 = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0 and "(3 of 7 chunks resumed, 900000000 evens)" in r2.stdout, r2.stdout
    assert open(path).read() == "".join(lines)


def test_one_writer_at_a_time(ctx, gb, tmp_path):
    path = str(tmp_path / "l.log")
    with open(path, "w") as f:
        fcntl.flock(f, fcntl.LOCK_EX)                    # another run holds the log
        with pytest.raises(gb.GoldbachError) as e:
            ctx.verify_range_logged(4, 10**6, 10**5, path, chunk_evens=10**5)
        assert e.value.status == gb.GB_EAGAIN and "in use" in e.value.message
    st, info = ctx.verify_range_logged(4, 10**6, 10**5, path, chunk_evens=10**5)
    assert info["complete"] == 1 and st["evens"] == (10**6 - 4) // 2 + 1
