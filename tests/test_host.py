"""CPU-side checks of the product's host pieces: the C ABI library exports
every symbol include/goldbach_gpu.h declares (no compute calls here), the C++
drop-in API (goldbach::PrimeBitset, simple_sieve, segmented_sieve) matches the
oracle word for word, and the CLI drivers keep the reference's argument
contract (tests/validation_gpu.sh:89-179, tests/validation.sh:56-95)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from _helpers import PKG_DIR, ROOT

HEADER = os.path.join(ROOT, "include", "goldbach_gpu.h")
BIN = os.path.join(PKG_DIR, "bin")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(gb_\w+)\s*\(", text, re.M)))


def test_abi_exports_every_declared_symbol(gb):
    lib = ctypes.CDLL(gb.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(gb.EXPORTS) == syms


def test_probe_library_exports_the_issue_probes():
    # bench.py's VALU and LDS ceilings come from this library (loaded only, no device call here)
    lib = ctypes.CDLL(os.path.join(PKG_DIR, "lib", "libgb_probe.so"))
    for s in ("gb_probe_valu", "gb_probe_lds"):
        assert hasattr(lib, s), s


def test_open_without_device_fails_loudly(gb):
    n = ctypes.c_int(0)
    if gb.library().gb_device_count(ctypes.byref(n)) == gb.GB_OK and n.value > 0:
        pytest.skip("a device is present")
    with pytest.raises(gb.GoldbachError) as e:
        gb.Context(0)
    assert e.value.status == gb.GB_ENODEV


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=120, **kw)


def test_known_answer_programs():
    for prog in ("test_sieve", "test_bitset"):
        r = run([os.path.join(BIN, prog)])
        assert r.returncode == 0 and "All tests passed." in r.stdout, r.stdout


@pytest.mark.parametrize("limit", [0, 2, 3, 4, 9, 30, 128, 129, 1000, 10**6 + 1, 10**7])
def test_cpp_prime_bitset_matches_oracle(orc, limit):
    r = run([os.path.join(BIN, "test_bitset"), "--dump", str(limit)])
    words = [int(x, 16) for x in r.stdout.split()]
    assert words == [int(x) for x in orc.prime_bitset(limit)]


@pytest.mark.parametrize("low,high", [(0, 1000), (1, 1), (999_983, 1_000_100), (10**10, 10**10 + 5000)])
def test_cpp_segmented_sieve_matches_oracle(orc, low, high):
    r = run([os.path.join(BIN, "test_sieve"), "--dump", str(low), str(high)])
    got = np.frombuffer(r.stdout.strip().encode(), dtype=np.uint8) - ord("0")
    assert np.array_equal(got, orc.segmented_sieve(low, high))


GOLDBACH = os.path.join(BIN, "goldbach")
SINGLE = os.path.join(BIN, "single_check")


@pytest.mark.parametrize("args,needle", [
    (["--help"], "Usage:"),
    (["--help"], "--seg-size=N"),
    (["1000000000", "--invalid"], "Error: Invalid"),
    (["--seg-size=1000000"], "Error: "),
    (["2", "--seg-size=1000000"], "Error: LIMIT must be >= 4"),
    (["1000000000", "--seg-size=1000001"], "Error: SEG_SIZE must be even"),
    (["1000000000", "--p-small=5000000000"], "Error: P_SMALL must be <= 4000000000"),
    (["1000000000", "--start=1000000002"], "Error: START must be less than LIMIT"),
    (["abc", "--seg-size=1000000"], "Error: Invalid numeric argument"),
])
def test_goldbach_cli_contract(args, needle):
    r = run([GOLDBACH] + args)
    assert needle in r.stdout + r.stderr
    if needle != "Usage:" and needle != "--seg-size=N":
        assert r.returncode == 1


@pytest.mark.parametrize("args,needle", [
    (["--help"], "Usage:"),
    (["100000000000000000000"], "Error: Number is too large"),
    (["1000000000001"], "Error: n must be an even integer"),
    (["2"], "Error: n must be >= 4"),
    (["abc"], "Error: Invalid numeric argument"),
])
def test_single_check_cli_contract(args, needle):
    r = run([SINGLE] + args)
    assert needle in r.stdout + r.stderr
