"""Host code under AddressSanitizer and UBSan (tests/sanitize/Makefile; no GPU involved):
the C++ host API's own test programs (PrimeBitset and segmented_sieve, the reference's
test_bitset.cpp / test_sieve.cpp checks restated), and the run-log reader fed intact logs
and 2,100 damaged ones (flipped and inserted bytes, truncations, deleted spans, 400-digit
numbers behind a valid checksum, 200 KB lines)."""
import os
import random
import shutil
import subprocess

import pytest

from test_runlog import header, line, record

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")

pytestmark = pytest.mark.skipif(not (shutil.which("g++") and os.path.exists("/opt/rocm/bin/hipcc")),
                                reason="needs g++ and hipcc")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-j4", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(SAN, "_build")


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, env=ENV)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
    return out


@pytest.mark.parametrize("prog", ["test_sieve", "test_bitset"])
def test_host_api_under_sanitizers(built, prog):
    assert "All tests passed." in run([os.path.join(built, prog)])


def mutations(seed=5, per_base=700):
    rng = random.Random(seed)
    base = [header() + "".join(record(i, sum_pmin=10 + i, max_pmin=3 + i, argmax=4 + 200 * i) for i in range(5)),
            header(chunk=7, pmin=1),
            header(start=10**18, limit=10**18 + 10**6, chunk=10**5)
            + "".join(record(i, start=10**18, limit=10**18 + 10**6, chunk=10**5) for i in range(3))]
    for b in base:
        yield b.encode()
        for _ in range(per_base):
            s = bytearray(b.encode())
            k = rng.randrange(6)
            if k == 0:
                for _ in range(rng.randrange(1, 6)):
                    s[rng.randrange(len(s))] = rng.randrange(256)
            elif k == 1:
                s = s[:rng.randrange(len(s) + 1)]
            elif k == 2:
                i = rng.randrange(len(s) + 1)
                s[i:i] = bytes(rng.randrange(256) for _ in range(rng.randrange(1, 40)))
            elif k == 3:
                s += line("chunk=0 first=" + "9" * rng.randrange(1, 400) + " last=4 evens=1").encode()
            elif k == 4:
                s += b"x" * rng.randrange(1, 200_000) + b"\n"
            else:
                i = rng.randrange(len(s) + 1)
                del s[i:i + rng.randrange(1, 50)]
            yield bytes(s)


def test_log_reader_under_sanitizers(built, tmp_path):
    paths = []
    for i, data in enumerate(mutations()):
        p = tmp_path / f"c{i:05d}.log"
        p.write_bytes(data)
        paths.append(str(p))
    out = run([os.path.join(built, "log_fuzz")] + paths)
    assert f"read {len(paths)} logs:" in out and "accepted" in out
