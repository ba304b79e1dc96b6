"""bin/cpu_goldbach, the product's drop-in for the reference's CPU oracle
(src/cpu_goldbach.cpp:45-130), on the product's segmented_sieve (libgoldbach.so).

Restates the reference's own contract rows for it (tests/validation.sh:35-37 the 1e6 run,
:56-66 --help, :73-83 the RAM guard at 1e14, a non-numeric LIMIT, LIMIT < 4) and compares
every line but the timings with the reference's cpu_goldbach built from its sources into
oracle/_ref."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "goldbach-gpu_amd", "bin")
CPU = os.path.join(BIN, "cpu_goldbach")
REF_CPU = os.path.join(ROOT, "oracle", "_ref", "cpu_goldbach")


def run(*args, binary=CPU, timeout=300):
    r = subprocess.run([binary, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def untimed(text):
    """the output without the lines that carry times"""
    return [ln for ln in text.splitlines() if not re.search(r"\bms\b", ln)]


def test_validation_1_1_oracle_1e6():
    rc, out = run(1000000)
    assert rc == 0 and "satisfy Goldbach. ✓" in out
    assert "Even numbers checked : 499999" in out and "Failures found       : 0" in out


@pytest.mark.parametrize("binary", ["cpu_goldbach", "big_check", "single_check"])
def test_validation_2_1_help(binary):
    rc, out = run("--help", binary=os.path.join(BIN, binary))
    assert rc == 0 and "Usage:" in out


def test_validation_3_1_ram_guard():
    rc, out = run(10**14)
    assert rc == 1 and "[!] ERROR: System RAM exhausted" in out


def test_validation_3_2_and_3_3_arguments():
    rc, out = run("abc")
    assert rc == 1 and "Error: Invalid numeric argument" in out
    rc, out = run(2)
    assert rc == 1 and "Error: LIMIT must be >= 4" in out
    # std::stoull's reading, as the reference's: an odd LIMIT is taken down to even
    rc, out = run(1001)
    assert rc == 0 and "All even numbers up to 1000 satisfy Goldbach. ✓" in out


@pytest.mark.parametrize("limit", [4, 5, 6, 100, 10**6, 10**7])
def test_same_output_as_the_reference(limit):
    if not os.path.exists(REF_CPU):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rc, out = run(limit)
    rc_ref, out_ref = run(limit, binary=REF_CPU)
    assert rc == rc_ref == 0
    assert untimed(out) == untimed(out_ref)
