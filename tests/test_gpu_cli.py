"""CLI contract on the device (tests/validation_gpu.sh:48-179 and
tests/validation.sh:39-41 of the reference), run against this build's drivers."""
import os
import subprocess

import pytest

from _helpers import PKG_DIR

pytestmark = pytest.mark.gpu
GOLDBACH = os.path.join(PKG_DIR, "bin", "goldbach")
SINGLE = os.path.join(PKG_DIR, "bin", "single_check")


def run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_basic_run(form):
    rc, out = run([GOLDBACH, "1000000000", "--seg-size=1000000", "--p-small=100000", "--batch-size=100000",
                   f"--form={form}"])
    assert rc == 0 and "satisfy Goldbach. ✓" in out and "Phase 2 fallbacks      : 0" in out


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_autotune_flag(form):
    rc, out = run([GOLDBACH, "100000000000", "--seg-size=134217728", f"--form={form}", "--autotune"])
    assert rc == 0 and "Autotuned      :" in out and "satisfy Goldbach. ✓" in out


def test_progress_flag():
    rc, out = run([GOLDBACH, "20000000000", "--seg-size=10000000", "--p-small=1000000", "--batch-size=1000000",
                   "--progress"])
    assert rc == 0 and "[Progress]" in out and "satisfy Goldbach. ✓" in out


def test_custom_start():
    rc, out = run([GOLDBACH, "1000000000", "--seg-size=1000000", "--p-small=100000", "--start=1000000"])
    assert rc == 0 and "Checking range : [1000000" in out and "satisfy Goldbach. ✓" in out


def test_vram_guard():
    # 288 GB of HBM holds the reference's 1e11-even segment; 1e13 evens does not fit
    rc, out = run([GOLDBACH, "100000000000", "--seg-size=10000000000000"])
    assert rc == 1 and "[!] ERROR: GPU" in out and "insufficient VRAM" in out


def test_more_gpus_than_available():
    rc, out = run([GOLDBACH, "1000000000", "--gpus=9"])
    assert "Requested" in out and "available" in out and "Using" in out


def test_forced_failure_reports_n():
    rc, out = run([GOLDBACH, "1000", "--p-small=3", "--phase2-limit=7"])
    assert rc == 1 and "[!] Goldbach FAILED at n = 98" in out


def test_single_check():
    rc, out = run([SINGLE, "1000000000000"])
    assert rc == 0 and "Goldbach holds. ✓" in out and "1000000000000 = 11 + 999999999989" in out


def test_counts_flag(orc):
    # c(n) lines in the reference's PRINT_COUNTS format (cpu_goldbach.cpp:101-103)
    rc, out = run([GOLDBACH, "10000", "--start=9000", "--counts"])
    assert rc == 0, out
    ref = orc.count_partitions(9000, 501)
    lines = [l for l in out.splitlines() if l.startswith("c(")]
    assert lines == [f"c({9000 + 2 * i}) = {int(c)}" for i, c in enumerate(ref)]
    i = int(ref.argmin())
    assert f"Smallest c(n)          : {int(ref[i])} at n = {9000 + 2 * i}" in out
