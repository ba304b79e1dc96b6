"""The run-log format read back without a device (gb_log_read): logs are written
here line by line with Python's zlib.crc32 (CRC-32/IEEE, an implementation
independent of the library's), then damaged the ways a crash or a disk can
damage them.  The reference resumes by hand with --start (goldbach.cu:581,
604-609, 689); the log replaces that, so what it accepts decides what a resumed
run skips."""
import zlib

import pytest


def line(body):
    return f"{body} crc={zlib.crc32(body.encode()):08x}\n"


def header(start=4, limit=1000, p_small=10**6, phase2_limit=10**8, chunk=100, pmin=0):
    return line(f"goldbach-run-log 1 start={start} limit={limit} p_small={p_small} "
                f"phase2_limit={phase2_limit} chunk_evens={chunk} pmin={pmin}")


def record(i, start=4, limit=1000, chunk=100, phase2=0, failures=0, first_fail=0, sum_pmin=0, max_pmin=0, argmax=0,
           first=None, last=None):
    a = start + 2 * i * chunk if first is None else first
    b = min(a + 2 * (chunk - 1), limit) if last is None else last
    return line(f"chunk={i} first={a} last={b} evens={(b - a) // 2 + 1} phase2={phase2} failures={failures} "
                f"first_fail={first_fail} sum_pmin={sum_pmin} max_pmin={max_pmin} argmax={argmax}")


def write(tmp_path, text, name="run.log"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_intact_log(gb, tmp_path):
    # [4, 1000] in chunks of 100 evens: 499 evens, the last chunk 99 of them
    recs = [record(i, sum_pmin=10 + i, max_pmin=3 + i, argmax=4 + 200 * i) for i in range(5)]
    h, got, torn = gb.read_log(write(tmp_path, header() + "".join(recs)))
    assert h == {"start": 4, "limit": 1000, "p_small": 10**6, "phase2_limit": 10**8, "chunk_evens": 100,
                 "want_pmin": 0}
    assert not torn and len(got) == 5
    assert got[4]["first_n"] == 804 and got[4]["last_n"] == 1000 and got[4]["evens"] == 99
    assert sum(r["evens"] for r in got) == 499
    assert [r["sum_pmin"] for r in got] == [10, 11, 12, 13, 14]


def test_header_only(gb, tmp_path):
    h, got, torn = gb.read_log(write(tmp_path, header(chunk=7, pmin=1)))
    assert h["chunk_evens"] == 7 and h["want_pmin"] == 1 and got == [] and not torn


def test_torn_last_line(gb, tmp_path):
    text = header() + record(0) + record(1) + record(2)[:30]
    _, got, torn = gb.read_log(write(tmp_path, text))
    assert len(got) == 2 and torn
    # a whole line without its newline is torn too
    _, got, torn = gb.read_log(write(tmp_path, header() + record(0) + record(1).rstrip("\n")))
    assert len(got) == 1 and torn


def test_corrupt_record_ends_the_prefix(gb, tmp_path):
    bad = record(1, phase2=5).replace("phase2=5", "phase2=6")      # checksum no longer matches
    _, got, torn = gb.read_log(write(tmp_path, header() + record(0) + bad + record(2)))
    assert len(got) == 1 and torn
    bad = record(1)[:-3] + "zz\n"                                   # checksum not hex
    _, got, _ = gb.read_log(write(tmp_path, header() + record(0) + bad))
    assert len(got) == 1


def test_records_must_follow_the_chunking(gb, tmp_path):
    # out of order, a repeated chunk, wrong bounds, a chunk past the range
    for recs in ([record(0), record(2)], [record(0), record(0)], [record(0), record(1, first=206)],
                 [record(0), record(1, last=400)], [record(i) for i in range(6)]):
        _, got, torn = gb.read_log(write(tmp_path, header() + "".join(recs)))
        assert torn and len(got) == len(recs) - 1, recs


def test_nothing_counts_after_a_failure(gb, tmp_path):
    text = header() + record(0) + record(1, failures=1, first_fail=300) + record(2)
    _, got, torn = gb.read_log(write(tmp_path, text))
    assert len(got) == 2 and got[1]["first_fail_n"] == 300 and torn


def test_non_canonical_lines_are_rejected(gb, tmp_path):
    rec = "chunk=0 first=004 last=202 evens=100 phase2=0 failures=0 first_fail=0 sum_pmin=0 max_pmin=0 argmax=0"
    _, got, torn = gb.read_log(write(tmp_path, header() + line(rec)))
    assert got == [] and torn
    rec = "chunk=0 first=4 last=202 evens=100 phase2=0 failures=0 first_fail=0 sum_pmin=0 max_pmin=0 argmax=0 x=1"
    _, got, _ = gb.read_log(write(tmp_path, header() + line(rec)))
    assert got == []


@pytest.mark.parametrize("text", [
    "hello\n",
    header().replace("limit=1000", "limit=1001"),                    # checksum broken
    line("goldbach-run-log 1 start=4 limit=1001 p_small=1 phase2_limit=1 chunk_evens=1 pmin=0"),   # odd limit
    line("goldbach-run-log 1 start=4 limit=1000 p_small=1 phase2_limit=1 chunk_evens=0 pmin=0"),   # no chunk
    line("goldbach-run-log 2 start=4 limit=1000 p_small=1 phase2_limit=1 chunk_evens=1 pmin=0"),   # version
    header().rstrip("\n"),                                          # torn header
])
def test_damaged_header_is_refused(gb, tmp_path, text):
    with pytest.raises(gb.GoldbachError) as e:
        gb.read_log(write(tmp_path, text))
    assert e.value.status == gb.GB_EINVAL


def test_missing_or_empty_log(gb, tmp_path):
    for path in (str(tmp_path / "absent.log"), write(tmp_path, "")):
        with pytest.raises(gb.GoldbachError) as e:
            gb.read_log(path)
        assert e.value.status == gb.GB_EIO
