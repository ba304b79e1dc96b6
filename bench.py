#!/usr/bin/env python3
"""Benchmark: even n verified per second over [4, 1e12] on one MI355X.

One step = one complete verification pass over all 499,999,999,999 even n in
[4, 1e12] with P_SMALL = 1e6, in the batched persistent form (one fused
sieve+mark kernel, two workgroups per CU, over mod-6 wheel tiles of 3 * 9 * 2^15 even n
from a device queue, in launches of at most 2^17 tiles, each replayed as a
hipGraph).  The plain per-segment form (2^27 evens per segment, as the reference's
loop) is timed on the same range beside it.

The JSON line carries:
  roofline      the fused wheel kernel.  Its binding resource is VALU issue: the
                wave-instructions one pass issues (rocprofv3 SQ_INSTS_VALU in the
                committed profile, used only when that profile was taken with the
                current kernel sources) over the pass's kernel time (HIP events on
                the library's stream around each launch), against the v_add_u32
                issue rate the VALU probe measures in the same run.  Beside it the
                HBM view: algorithmic bytes (0.375 B per even n: window bitset
                written + read, verified bitmap written) over the same time, and
                the measured traffic.  Without a current profile the HBM view is
                the roofline.  roofline.lds prices the same pass's counted
                SQ_INSTS_LDS (wave-level LDS instructions) against the LDS probe.
  step_spread   min / median / max / mean / stdev / CV of the timed passes
                (the protocol of benchmarks/measure_variance.sh:11-28).
  valu_probe    issue rates of independent v_add_u32, independent v_mad_u64_u32
                and a dependent v_add_u32 chain (lib/libgb_probe.so).
  lds_probe     wave-level LDS instruction rates: ds_and_b32 at scattered and at
                bank-disjoint words, ds_read_b32, ds_and_b32 from half the lanes; the
                scattered ds_and_b32 rate is the ceiling of roofline.lds (the wheel
                kernel's counted SQ_INSTS_LDS over its live time).
  plain         the plain form over [4, 1e12] (two segments in flight on two streams)
                with its host overhead; beside it
                [4, 1e11] in 2^27-even segments at P_SMALL = 1e6 and 4e9 (the cost
                must not grow with P_SMALL), with the sieve and marking kernels
                timed apart and priced against the HBM roofline, and [4, 1e9] in
                both forms, each against the reference's published RTX 5090 time,
                next to the product's cpu_goldbach and the reference's own
                cpu_goldbach 1000000000 (one thread, built into oracle/_ref; skip
                it with --no-ref-cpu).
  cpu_baseline  the CPU restatement (oracle/, OpenMP over segments, the
                reference algorithm: per-n ascending prime scan) timed on this
                host over 64 sampled 2^24-even segments spread over [4, 1e12].
  high_windows  64 windows above 1e15 (q by pre-sieve + Baillie-PSW, the survivors of one
                window word in five proven, the n left open decided by their other
                survivors, then exactly by the slow path) in persistent launches, evens/s,
                with its counted VALU issue against
                the probe's ceiling; its CPU baseline is the oracle's
                Miller-Rabin restatement on the first 2^19 evens of each window.
  partition_counts  c(n) for every even n in [4, 3e7] (gb_count_partitions) in
                candidate pairs/s (the reference loop's p = 2..n/2 per n), with
                the oracle's c(n) on 512 n below 3e7 as its CPU baseline.
  big_check     gb_big_min_partition for 10^1000 and 10^10000 (seconds per n,
                against the reference's published 20-thread GMP times), with GMP's
                own test over the same primes as its CPU baseline.
  serving       the continuous-batching server (gb_serve_*): p50/p99 latency
                from scheduled arrival to result under open-loop Poisson
                arrivals at 1e4, 1e5 and 1e6 queries/s, and queries/s for a
                2^20-query burst; n log-uniform in [1e6, 2^64-2]; its CPU
                baseline is the oracle's min_partition over 2^21 queries of the stream.
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import math
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

LIMIT = 10**12
P_SMALL = 10**6
PLAIN_SEG = 1 << 27
HBM_PEAK_GBS = 8000.0
CUS = 256                      # MI355X compute units
CLOCK_HZ = 2.4e9               # MI355X peak engine clock
ALG_BYTES_PER_EVEN = 0.375
REF_PUBLISHED_EVENS_PER_S = 499_999_999_999 / 37.440   # RESULTS.md:48, 1x RTX 5090
METRIC = "even n verified/s at N=1e12 on 1 GPU, batched vs plain; p99 query latency"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "wheel_kernel_1e12_traffic.json")
WINDOWS_PROFILE = os.path.join(ROOT, "profiles", "high_windows_mr.json")
REF_CPU_GOLDBACH = os.path.join(ROOT, "oracle", "_ref", "cpu_goldbach")   # the reference's CPU program, built from its sources
REF_GPU_SECONDS = {10**9: 0.141018, 10**11: 3.3115}   # RESULTS.md:45,47, 1x RTX 5090
PRESIEVE = 1 << 22             # the high-window tiles pre-sieve with every table prime below 2^22
SQR_VALU = 28                  # VALU instructions of one 64-bit Montgomery square (mont_sqr, gfx950)
PRODUCT_CPU_GOLDBACH = os.path.join(ROOT, "goldbach-gpu_amd", "bin", "cpu_goldbach")
PROBE_LIB = os.path.join(ROOT, "goldbach-gpu_amd", "lib", "libgb_probe.so")
# sources whose change makes a committed counter profile stale: the kernels and their launch geometry
PROFILED_SOURCES = [os.path.join(ROOT, "goldbach-gpu_amd", "csrc", f)
                    for f in ("gb_kernels.hip", "gb_device.hpp", "gb_kernels.hpp", "gb_api.cpp", "gb_ctx.hpp")]

# The rocprofv3 runs behind the committed counter profiles: name -> (rocprofv3 options, program).
# `bench.py --profile-commands DIR` prints them (each under its own time limit), and
# `bench.py --summarize-profiles DIR` rebuilds profiles/*.json from their CSV output.
HEADLINE_CMD = ("python3 -u bench.py --steps 1 --warmup 0 --no-cpu --no-ref-cpu --no-plain --no-serve --no-windows "
                "--no-counts --no-big")
WINDOWS_CMD = "python3 -u bench.py --steps 1 --warmup 0 --no-cpu --no-ref-cpu --no-plain --no-serve --no-counts --no-big"
PROFILE_RUNS = {
    "t": ("--kernel-trace --stats", HEADLINE_CMD),
    "s": ("--pmc SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_WAVES SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE "
          "SQ_BUSY_CYCLES SQ_WAVE_CYCLES", HEADLINE_CMD),
    "a": ("--pmc SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY SQC_ICACHE_MISSES", HEADLINE_CMD),
    "f": ("--pmc FETCH_SIZE", HEADLINE_CMD),
    "w": ("--pmc WRITE_SIZE", HEADLINE_CMD),
    "cf": ("--pmc FETCH_SIZE", "python3 -u bench.py --copy-probe"),
    "cw": ("--pmc WRITE_SIZE", "python3 -u bench.py --copy-probe"),
    "ht": ("--kernel-trace --stats", WINDOWS_CMD),
    "hs": ("--pmc SQ_INSTS_VALU SQ_INSTS_VALU_INT32 SQ_INSTS_VALU_INT64 SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVES "
           "SQ_BUSY_CYCLES", WINDOWS_CMD),
    "p": ("--pmc SQ_INSTS_VALU SQ_BUSY_CYCLES SQ_WAVES", "python3 -u bench.py --valu-probe"),
    "l": ("--pmc SQ_INSTS_LDS SQ_INSTS_VALU SQ_LDS_IDX_ACTIVE SQ_LDS_BANK_CONFLICT SQ_BUSY_CYCLES SQ_WAVES",
          "python3 -u bench.py --lds-probe"),
}
COPY_PROBE_BYTES = 1 << 30


def profile_command(name: str, out_dir: str) -> str:
    opts, prog = PROFILE_RUNS[name]
    return f"rocprofv3 {opts} -f csv -d {out_dir}/{name} -o {name} -- {prog}"


def _csv_rows(out_dir: str, name: str, suffix: str) -> list:
    import csv
    import glob

    rows = []
    for path in sorted(glob.glob(os.path.join(out_dir, name, "**", f"*{suffix}.csv"), recursive=True)):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no {suffix}.csv under {os.path.join(out_dir, name)}")
    return rows


def _dispatches(out_dir: str, name: str) -> list:
    """Counter values per kernel dispatch, in dispatch order: [(kernel name, {counter: value}, ns)]."""
    by = {}
    for r in _csv_rows(out_dir, name, "counter_collection"):
        d = by.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], {}, 0])
        d[1][r["Counter_Name"]] = d[1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        d[2] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return [tuple(by[k]) for k in sorted(by)]


def _kernel_stats(out_dir: str, name: str, key: str) -> dict:
    for r in _csv_rows(out_dir, name, "kernel_stats"):
        if key in r["Name"]:
            return {"calls": int(r["Calls"]), "total_ms": float(r["TotalDurationNs"]) * 1e-6,
                    "avg_ms": float(r["AverageNs"]) * 1e-6, "name": r["Name"].split("(")[0].replace("void ", "")}
    raise SystemExit(f"{key} not in the kernel stats of {name}")


def summarize_profiles(out_dir: str) -> int:
    """Rebuild profiles/wheel_kernel_1e12_traffic.json and profiles/high_windows_mr.json from the
    CSV output of the PROFILE_RUNS under out_dir, stamped with the current build hash."""
    wheel, fused = "wheel_kernel<512", "fused_kernel<1024>"

    def summed(name, key):
        tot = {}
        for kname, ctr, _ in _dispatches(out_dir, name):
            if key in kname:
                for c, v in ctr.items():
                    tot[c] = tot.get(c, 0.0) + v
        return tot

    def copy_kib(name, counter):
        vals = [ctr[counter] for kname, ctr, _ in _dispatches(out_dir, name) if "copyBuffer" in kname]
        return max(vals)

    cf_kib, cw_kib = copy_kib("cf", "FETCH_SIZE"), copy_kib("cw", "WRITE_SIZE")
    fetch_scale = round(COPY_PROBE_BYTES / (cf_kib * 1024.0), 3)
    write_scale = round(COPY_PROBE_BYTES / (cw_kib * 1024.0), 3)
    trace = _kernel_stats(out_dir, "t", wheel)
    s, a = summed("s", wheel), summed("a", wheel)
    fetch, write = summed("f", wheel)["FETCH_SIZE"], summed("w", wheel)["WRITE_SIZE"]
    hbm = fetch * 1024 * fetch_scale + write * 1024 * write_scale
    prof = {
        "kernel": trace["name"],
        "workload": "all even n in [4, 1e12], batched persistent form, mod-6 wheel tiles of 3 x 9 x 2^15 evens, two "
                    "workgroups per CU, P_SMALL = 1e6: one pass = %d launches of at most 2^17 tiles" % trace["calls"],
        "build_hash": build_hash(),
        "launches_per_pass": trace["calls"],
        "trace_ms_per_pass": round(trace["total_ms"], 3),
        "trace_ms_per_launch": round(trace["avg_ms"], 3),
        "fetch_size_kib_per_pass": fetch, "fetch_scale": fetch_scale,
        "write_size_kib_per_pass": write, "write_scale": write_scale,
        "hbm_bytes_per_pass": int(hbm),
        "sq_insts_valu_per_pass": s["SQ_INSTS_VALU"], "sq_insts_lds_per_pass": s["SQ_INSTS_LDS"],
        "sq_insts_salu_per_pass": s["SQ_INSTS_SALU"], "sq_lds_bank_conflict": s["SQ_LDS_BANK_CONFLICT"],
        "sq_lds_idx_active": s["SQ_LDS_IDX_ACTIVE"], "sq_busy_cycles": s["SQ_BUSY_CYCLES"],
        "sq_wave_cycles": s["SQ_WAVE_CYCLES"], "sq_waves": s["SQ_WAVES"],
        "wait": {k: a[k] for k in sorted(a)},
        "calibration": "1 GiB device-to-device copy (bench.py --copy-probe): FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB"
                       % (cf_kib, cw_kib),
        "commands": [profile_command(n, "<dir>") for n in ("t", "s", "a", "f", "w", "cf", "cw")],
        "note": "counter values summed over the pass's wheel_kernel dispatches (rocprofv3 CSV output, "
                "counter_collection rows per dispatch)",
    }
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(prof, f, indent=1)
        f.write("\n")
    runs = [(ctr, ns) for kname, ctr, ns in _dispatches(out_dir, "hs") if fused in kname]
    hs_all = summed("hs", fused)
    htrace = _kernel_stats(out_dir, "ht", fused)
    # dispatch order: the warm-up call (the windows cut to 2^20 evens: one launch), then the measured
    # call's launches
    main_ctr = {c: sum(ctr[c] for ctr, _ in runs[1:]) for c in runs[0][0]}
    main_ns = sum(ns for _, ns in runs[1:])
    hprof = {
        "kernel": "gbk::fused_kernel<1024>",
        "workload": "64 windows of uniform 1e6..1e9 evens at even starts log-uniform in [1e15, 1.8e19] (numpy seed "
                    "0): pre-sieve with every odd prime below 2^22, base-2 strong test, strong Lucas test",
        "build_hash": build_hash(),
        "measured_launches": len(runs) - 1,
        "sq_insts_valu_per_call": main_ctr["SQ_INSTS_VALU"],
        "sq_insts_valu_int64_per_call": main_ctr["SQ_INSTS_VALU_INT64"],
        "sq_insts_valu_int32_per_call": main_ctr["SQ_INSTS_VALU_INT32"],
        "kernel_ms_under_profiler": round(main_ns * 1e-6, 3),
        "trace_launches": htrace["calls"], "trace_ms_all_launches": round(htrace["total_ms"], 3),
        "sq_insts_valu_int64_launches": hs_all["SQ_INSTS_VALU_INT64"],
        "sq_insts_valu_int32_launches": hs_all["SQ_INSTS_VALU_INT32"],
        "sq_insts_salu_launches": hs_all["SQ_INSTS_SALU"], "sq_insts_lds_launches": hs_all["SQ_INSTS_LDS"],
        "commands": [profile_command(n, "<dir>") for n in ("ht", "hs")],
        "note": "the *_per_call figures sum the measured call's fused_kernel dispatches (the first dispatch is the "
                "warm-up on the same windows cut to 2^20 evens); the *_launches figures sum all of them",
    }
    with open(WINDOWS_PROFILE, "w") as f:
        json.dump(hprof, f, indent=1)
        f.write("\n")
    probe = [(kname[kname.index("valu_probe"):kname.index(">") + 1], ctr, ns)
             for kname, ctr, ns in _dispatches(out_dir, "p") if "valu_probe" in kname]
    lprobe = []
    if os.path.isdir(os.path.join(out_dir, "l")):
        lprobe = [(kname[kname.index("lds_probe"):kname.index(">") + 1], ctr, ns)
                  for kname, ctr, ns in _dispatches(out_dir, "l") if "lds_probe" in kname]
    print(json.dumps({"wheel": prof, "windows": hprof, "windows_trace": htrace, "copy_kib": [cf_kib, cw_kib],
                      "probe": [{"kernel": k, "ns": ns, **c} for k, c, ns in probe],
                      "lds_probe": [{"kernel": k, "ns": ns, **c} for k, c, ns in lprobe]}, indent=1))
    return 0


def build_hash() -> str:
    """16 hex digits of SHA-256 over the profiled sources; committed profiles carry the value
    they were taken with, and a figure read from a profile of other sources is reported as null."""
    h = hashlib.sha256()
    for f in PROFILED_SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def load_profile(path: str):
    """A committed counter profile if it was taken with the current sources, else None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        prof = json.load(f)
    return prof if prof.get("build_hash") == build_hash() else None


def valu_probe(kinds=(0, 1, 2), blocks: int = 256 * 64, iters: int = 2048) -> dict:
    """Achievable VALU issue, measured: gb_probe_valu runs fixed counts of v_add_u32 (kind 0,
    eight independent chains), v_mad_u64_u32 (kind 1) and one dependent v_add_u32 chain (kind 2)
    over 16,384 workgroups; rate = wave-instructions issued / launch time (HIP events)."""
    lib = ctypes.CDLL(PROBE_LIB)
    lib.gb_probe_valu.restype = ctypes.c_int
    lib.gb_probe_valu.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    names = {0: "v_add_u32_independent", 1: "v_mad_u64_u32_independent", 2: "v_add_u32_dependent"}
    out = {}
    for k in kinds:
        best = None
        for _ in range(3):                                  # the first launch loads the code
            ms, insts = ctypes.c_double(), ctypes.c_double()
            rc = lib.gb_probe_valu(k, blocks, iters, ctypes.byref(ms), ctypes.byref(insts))
            if rc != 0:
                raise RuntimeError(f"gb_probe_valu({k}) failed: {rc}")
            rate = insts.value / (ms.value * 1e-3)
            best = rate if best is None else max(best, rate)
        out[names[k]] = {"wave_insts_per_s": best, "per_cu_clock_at_2.4GHz": best / (CUS * CLOCK_HZ)}
    return out


def lds_probe(kinds=(0, 1, 2, 3), iters: int = 2048) -> dict:
    """Achievable LDS issue, measured: gb_probe_lds runs fixed counts of 32-bit LDS operations
    from two 512-thread workgroups per CU (ds_and_b32 at scattered words, at words 64x + lane,
    ds_read_b32 at scattered words, ds_and_b32 from half the lanes); rate = wave-level LDS
    instructions / launch time (HIP events).  The scattered and the bank-disjoint addresses run
    at nearly one rate, and half an exec mask saves little: the LDS cost is per wave-level
    instruction, about six CU clocks each."""
    lib = ctypes.CDLL(PROBE_LIB)
    lib.gb_probe_lds.restype = ctypes.c_int
    lib.gb_probe_lds.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    names = {0: "ds_and_b32_scattered", 1: "ds_and_b32_bank_disjoint", 2: "ds_read_b32_scattered",
             3: "ds_and_b32_half_exec"}
    out = {}
    for k in kinds:
        best = None
        for _ in range(3):                                  # the first launch loads the code
            ms, insts = ctypes.c_double(), ctypes.c_double()
            rc = lib.gb_probe_lds(k, CUS * 2, iters, ctypes.byref(ms), ctypes.byref(insts))
            if rc != 0:
                raise RuntimeError(f"gb_probe_lds({k}) failed: {rc}")
            rate = insts.value / (ms.value * 1e-3)
            best = rate if best is None else max(best, rate)
        out[names[k]] = {"wave_insts_per_s": best, "cu_clocks_per_inst_at_2.4GHz": CUS * CLOCK_HZ / best}
    return out


def device_sync() -> None:
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def timed_passes(fn, steps: int, warmup: int):
    """(total seconds, outputs, per-step seconds) of `steps` passes after `warmup` untimed ones."""
    for _ in range(warmup):
        fn()
    device_sync()
    t0 = time.perf_counter()
    outs, per = [], []
    for _ in range(steps):
        t = time.perf_counter()
        outs.append(fn())          # each pass returns after its last launch finished
        per.append(time.perf_counter() - t)
    device_sync()
    return time.perf_counter() - t0, outs, per


def spread(per_s) -> dict:
    """Run-to-run spread of the timed passes (benchmarks/measure_variance.sh:11-28 reports
    mean, standard deviation and coefficient of variation over 20 runs after a warm-up)."""
    ms = [x * 1e3 for x in per_s]
    mean = statistics.fmean(ms)
    sd = statistics.stdev(ms) if len(ms) > 1 else 0.0
    return {"runs": len(ms), "min_ms": min(ms), "median_ms": statistics.median(ms), "max_ms": max(ms),
            "mean_ms": mean, "stdev_ms": sd, "cv": sd / mean if mean else None}


def host_cpu() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline() -> dict:
    from _helpers import Oracle

    orc = Oracle()
    threads = min(16, os.cpu_count() or 1)
    seg = 1 << 24
    starts = [4 + 2 * ((k * (LIMIT // 2 - seg)) // 63) for k in range(64)]
    evens = 0
    t0 = time.perf_counter()
    from concurrent.futures import ThreadPoolExecutor

    def one(a):
        st, _, _ = orc.verify_segment(a, seg, P_SMALL, want_bits=False, want_pmin=False)
        return st["evens"], st["failures"]

    with ThreadPoolExecutor(max_workers=threads) as ex:
        for e, f in ex.map(one, starts):
            evens += e
            assert f == 0
    dt = time.perf_counter() - t0
    return {"value": evens / dt, "unit": "evens/s", "cores": threads, "kind": "port",
            "sample": f"64 segments of 2^24 evens spread over [4, 1e12], {threads} threads, {dt:.1f} s wall"}


def cpu_windows_baseline() -> dict:
    """The oracle (u128 Miller-Rabin restatement) on the first 2^19 evens of each of the 64 high windows."""
    from concurrent.futures import ThreadPoolExecutor

    from _helpers import Oracle

    orc = Oracle()
    threads = min(16, os.cpu_count() or 1)
    seg = 1 << 19
    t0 = time.perf_counter()

    def one(w):
        st, _, _ = orc.verify_segment(w[0], seg, P_SMALL, want_bits=False, want_pmin=False)
        assert st["failures"] == 0
        return st["evens"]

    with ThreadPoolExecutor(max_workers=threads) as ex:
        evens = sum(ex.map(one, high_windows()))
    dt = time.perf_counter() - t0
    return {"value": evens / dt, "unit": "evens/s", "cores": threads, "kind": "port",
            "sample": f"the first 2^19 evens of each of the 64 windows, {threads} threads, {dt:.1f} s wall"}


def cpu_serving_baseline() -> dict:
    """The oracle's min_partition over 2^21 queries of the serving stream (OpenMP over queries)."""
    from _helpers import Oracle

    orc = Oracle()
    threads = min(16, os.cpu_count() or 1)
    ns = serving_ns(1 << 21, seed=1)
    t0 = time.perf_counter()
    p = orc.min_partition_batch(ns, threads=threads)
    dt = time.perf_counter() - t0
    assert int((p == 0).sum()) == 0
    return {"value": ns.shape[0] / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"2^21 queries of the serving stream (n log-uniform in [1e6, 2^64-2], seed 1), "
                      f"{threads} threads, {dt:.1f} s wall"}


def serving_ns(count: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(1e6), np.log(2.0**64 - 2), size=count))
    return np.maximum((x / 2).astype(np.uint64) * np.uint64(2), np.uint64(4))


def poisson_level(srv, ns: np.ndarray, qps: float, seconds: float, seed: int) -> dict:
    """Open-loop Poisson arrivals at `qps`: latency = completion - scheduled arrival."""
    rng = np.random.default_rng(seed)
    count = max(1, min(ns.shape[0], int(qps * seconds)))
    gaps = rng.exponential(1e9 / qps, size=count)
    t0 = time.monotonic_ns() + 2_000_000
    arrive = t0 + np.cumsum(gaps).astype(np.int64)
    first = None
    i = 0
    parts = []
    done = 0
    while i < count:
        now = time.monotonic_ns()
        j = int(np.searchsorted(arrive, now, side="right"))
        if j > i:
            fid = srv.submit(ns[i:j])
            first = fid if first is None else first
            i = j
        r = srv.poll(65536)
        if r.shape[0]:
            parts.append(r)
            done += r.shape[0]
    while done < count:
        r = srv.poll(65536, wait_ms=20)
        parts.append(r)
        done += r.shape[0]
    r = np.concatenate(parts)
    r = r[np.argsort(r["id"])]
    assert r.shape[0] == count and int(r["id"][0]) == first
    lat_ms = (r["done_ns"].astype(np.int64) - arrive) / 1e6
    span = (int(r["done_ns"].max()) - int(arrive[0])) / 1e9
    return {"offered_qps": qps, "queries": count, "p50_ms": float(np.percentile(lat_ms, 50)),
            "p99_ms": float(np.percentile(lat_ms, 99)), "achieved_qps": count / span,
            "mean_steps": float(r["steps"].mean()), "p": r["p"]}


def serving(ctx) -> dict:
    """Continuous-batching server: Poisson arrivals at fixed rates, then a saturating burst."""
    ns = serving_ns(1 << 20, seed=1)
    ref = ctx.min_partition(ns)          # batch answers: every served p must equal them
    levels = []
    with ctx.server() as srv:
        srv.run(ns)                      # warm: a server in service holds a burst's result blocks
        for k, qps in enumerate((1e4, 1e5, 1e6)):
            lv = poisson_level(srv, ns, qps, 1.0, seed=100 + k)
            assert np.array_equal(lv.pop("p"), ref[:lv["queries"]])
            levels.append(lv)
        t0 = time.monotonic_ns()
        first = srv.submit(ns)
        parts, done = [], 0
        while done < ns.shape[0]:
            r = srv.poll(65536, wait_ms=20)
            parts.append(r)
            done += r.shape[0]
        r = np.concatenate(parts)
        r = r[np.argsort(r["id"])]
        assert int(r["id"][0]) == first and np.array_equal(r["p"], ref)
        burst_s = (int(r["done_ns"].max()) - t0) / 1e9
        st = srv.stats()
    return {"levels": levels, "p50_ms": levels[0]["p50_ms"], "p99_ms": levels[0]["p99_ms"],
            "burst_qps": ns.shape[0] / burst_s, "burst": int(ns.shape[0]), "slots": st["slots"],
            "mode": "continuous batching, open-loop Poisson arrivals (seed 100+k) of n log-uniform in "
                    "[1e6, 2^64-2] (seed 1); headline p50/p99 at the lowest rate"}


def high_windows(count: int = 64, lo_evens: int = 10**6, hi_evens: int = 10**9, seed: int = 0):
    """64 windows of uneven length (uniform 1e6..1e9 evens) at even starts log-uniform in [1e15, 1.8e19]."""
    rng = np.random.default_rng(seed)
    starts = np.exp(rng.uniform(np.log(1e15), np.log(1.8e19), size=count))
    lens = rng.integers(lo_evens, hi_evens + 1, size=count)
    return [(int(a / 2) * 2, int(e)) for a, e in zip(starts, lens)]


def plain_1e11(ctx, gb) -> dict:
    """The plain form on [4, 1e11] in 2^27-even segments (the reference's published setting,
    goldbach 100000000000 --seg-size=134217728) at P_SMALL = 1e6 and at the allowed maximum 4e9:
    the answers agree and the cost does not grow with P_SMALL.  A third pass on a context with
    per-kernel timing splits the device time between the sieve and the marking kernels and
    prices each against the HBM roofline with its algorithmic bytes: the sieve writes one bit
    per odd q of its window and reads a 16-byte entry per sieving prime; the marking reads the
    window bits and writes one verified bit per even n."""
    limit, evens = 10**11, (10**11 - 4) // 2 + 1
    out = {"workload": "all even n in [4, 1e11], plain form, 2^27 evens per segment",
           "reference_gpu_seconds": REF_GPU_SECONDS[limit]}
    ctx.verify_range(4, limit, P_SMALL, form=gb.GB_FORM_PLAIN, seg_evens=PLAIN_SEG)   # warm
    ref = None
    for ps in (P_SMALL, 4 * 10**9):
        device_sync()
        t0 = time.perf_counter()
        st = ctx.verify_range(4, limit, ps, form=gb.GB_FORM_PLAIN, seg_evens=PLAIN_SEG)
        device_sync()
        dt = time.perf_counter() - t0
        assert st["evens"] == evens and st["failures"] == 0 and st["phase2"] == 0
        ref = ref or st
        assert all(st[k] == ref[k] for k in ("evens", "phase2", "failures"))
        out[f"p_small_{ps:.0e}".replace("+", "")] = {"seconds": dt, "evens_per_s": evens / dt,
                                                     "kernel_ms": st["kernel_ms"],
                                                     "host_overhead": dt * 1e3 / st["kernel_ms"] - 1.0,
                                                     "vs_reference": REF_GPU_SECONDS[limit] / dt}
    with gb.Context(0, max_seg_evens=PLAIN_SEG, kernel_times=True) as kc:
        kc.verify_range(4, 10**9, P_SMALL, form=gb.GB_FORM_PLAIN, seg_evens=PLAIN_SEG)   # warm
        kc.kernel_times(reset=True)
        st = kc.verify_range(4, limit, P_SMALL, form=gb.GB_FORM_PLAIN, seg_evens=PLAIN_SEG)
        kt = kc.kernel_times(reset=True)
    assert st["evens"] == evens and kt["evens"] == evens
    sieve_bytes = kt["window_bits"] / 8 + kt["sieve_primes"] * 16
    mark_bytes = kt["window_bits"] / 8 + kt["evens"] / 8
    out["kernels"] = {
        "segments": kt["segments"],
        "sieve_window_kernel": {"ms": kt["sieve_ms"], "alg_bytes": sieve_bytes,
                                "achieved_gbs": sieve_bytes / (kt["sieve_ms"] * 1e-3) / 1e9,
                                "hbm_frac": sieve_bytes / (kt["sieve_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "odds_per_s": kt["window_bits"] / (kt["sieve_ms"] * 1e-3)},
        "mark_segment_kernel": {"ms": kt["mark_ms"], "alg_bytes": mark_bytes,
                                "achieved_gbs": mark_bytes / (kt["mark_ms"] * 1e-3) / 1e9,
                                "hbm_frac": mark_bytes / (kt["mark_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "other_ms": kt["other_ms"],
        "note": "HIP events between the launches of every segment (a context opened with "
                "GB_CFG_KERNEL_TIMES); peak 8 TB/s"}
    return out


def cpu_goldbach_run(path: str, limit: int) -> dict:
    t0 = time.perf_counter()
    r = subprocess.run([path, str(limit)], capture_output=True, text=True, timeout=900)
    dt = time.perf_counter() - t0
    assert r.returncode == 0 and "satisfy Goldbach" in r.stdout, r.stdout + r.stderr
    evens = (limit - 4) // 2 + 1
    assert f"Even numbers checked : {evens}" in r.stdout
    return {"seconds": dt, "evens_per_s": evens / dt}


def limit_1e9(ctx, gb, ref_cpu: bool) -> dict:
    """All even n in [4, 1e9] in the plain and batched forms; beside them the CPU programs over
    the same range: the product's cpu_goldbach (OpenMP, on the product's segmented_sieve) and,
    with ref_cpu, the reference's own cpu_goldbach 1000000000 (one thread, built from its
    sources into oracle/_ref), the CPU baseline of this workload."""
    limit, evens = 10**9, (10**9 - 4) // 2 + 1
    out = {"workload": "all even n in [4, 1e9]", "reference_gpu_seconds": REF_GPU_SECONDS[limit]}
    for name, kw in (("plain", {"form": gb.GB_FORM_PLAIN, "seg_evens": PLAIN_SEG}),
                     ("batched", {"form": gb.GB_FORM_BATCHED})):
        ctx.verify_range(4, limit, P_SMALL, **kw)                                   # warm
        device_sync()
        t0 = time.perf_counter()
        st = ctx.verify_range(4, limit, P_SMALL, **kw)
        device_sync()
        dt = time.perf_counter() - t0
        assert st["evens"] == evens and st["failures"] == 0 and st["phase2"] == 0
        out[name] = {"seconds": dt, "evens_per_s": evens / dt, "vs_reference": REF_GPU_SECONDS[limit] / dt}
    cpu = host_cpu()
    if os.path.exists(PRODUCT_CPU_GOLDBACH):
        out["product_cpu"] = {**cpu_goldbach_run(PRODUCT_CPU_GOLDBACH, limit), "threads": cpu["affinity"],
                              "program": "goldbach-gpu_amd/bin/cpu_goldbach 1000000000", "host": cpu}
    if ref_cpu and os.path.exists(REF_CPU_GOLDBACH):
        out["reference_cpu"] = {**cpu_goldbach_run(REF_CPU_GOLDBACH, limit), "threads": 1,
                                "program": "oracle/_ref/cpu_goldbach 1000000000 (the reference's "
                                           "cpu_goldbach.cpp + segmented_sieve.cpp, g++ -O3)", "host": cpu}
    return out


def presieve_density() -> float:
    """Fraction of odd q with no prime factor in [3, 2^22): the product of (1 - 1/p) over the table."""
    s = np.ones(PRESIEVE, dtype=bool)
    s[:2] = False
    for i in range(2, int(PRESIEVE ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    ps = np.nonzero(s)[0][1:].astype(np.float64)          # odd primes
    return float(np.prod(1.0 - 1.0 / ps))


MR_PROOF_STRIDE = 5   # the tiles prove the survivors of one window word in five (gb_tuning default)


def montgomery_products(wins, stride: int = MR_PROOF_STRIDE) -> float:
    """Model of the 64-bit Montgomery products the windows need (one product = one mulmod).

    A window of e evens near q holds about e odd q; e * S of them survive the pre-sieve
    (S = presieve_density()), and the tiles test those in one window word in `stride`: the
    base-2 strong test, b - 2 products for q of b bits (b - 3 squarings of the ladder over d,
    n - 1 = d 2^r with E[r] = 2, and E[r - 1] = 1 more); the primes among them (2/ln q of the odd
    q) also take the strong Lucas test, 2b - 7 products (two per bit of m = (d - 1)/2,
    n + 1 = d 2^s, one for the top bit, one for W_d, E[s - 1] = 1).  The n no proven q
    resolves (about 3e-4 of them at stride 5) take a few more tests each, which the model
    leaves out (under 1 % of the products).
    """
    dens = presieve_density()
    total = 0.0
    for a, e in wins:
        q = a + e
        b = q.bit_length()
        total += (e * dens * (b - 2) + e * 2.0 / math.log(q) * (2 * b - 7)) / stride
    return total


def windows_leg(ctx, probe) -> dict:
    """High windows: q decided by pre-sieve + Baillie-PSW, all windows in persistent launches of
    at most 2^15 tiles.
    Its roofline is VALU issue: the VALU wave-instructions the launch issues (rocprofv3 count
    in the committed profile, when that profile was taken with the current sources) over the
    live launch time, against the v_add_u32 issue rate the probe measures."""
    wins = high_windows()
    evens = sum(e for _, e in wins)
    ctx.verify_windows([(a, min(e, 1 << 20)) for a, e in wins], P_SMALL)   # warm (tables, code)
    device_sync()
    t0 = time.perf_counter()
    per, tot = ctx.verify_windows(wins, P_SMALL)
    device_sync()
    dt = time.perf_counter() - t0
    assert tot["evens"] == evens and tot["failures"] == 0
    products = montgomery_products(wins)
    ksec = tot["kernel_ms"] * 1e-3
    prof = load_profile(WINDOWS_PROFILE)
    peak = probe["v_add_u32_independent"]["wave_insts_per_s"] if probe else None
    achieved = prof["sq_insts_valu_per_call"] / ksec if prof else None
    # 64-bit integer instructions (v_mad_u64_u32, 64-bit adds) issue at half the v_add_u32 rate:
    # counted as two issue slots each, the launch's slots over the live time against the same peak
    int64 = prof.get("sq_insts_valu_int64_per_call") if prof else None
    slots = (prof["sq_insts_valu_per_call"] + int64) / ksec if int64 else None
    return {"workload": "64 windows, uniform 1e6..1e9 evens, starts log-uniform in [1e15, 1.8e19] (seed 0)",
            "evens": evens, "value": evens / dt, "unit": "evens/s", "seconds": dt,
            "kernel_ms": tot["kernel_ms"], "phase2": tot["phase2"],
            "roofline": {"bound": "valu", "unit": "wave-instructions/s", "achieved": achieved, "peak": peak,
                         "frac": achieved / peak if (achieved and peak) else None,
                         "kernel": "fused_kernel<1024>",
                         "counted": "SQ_INSTS_VALU of the call's launches, profiles/high_windows_mr.json (null when "
                                    "that profile was taken with other sources)",
                         "peak_note": "v_add_u32 issue rate of gb_probe_valu, measured in this run",
                         "issue_slots_per_s": slots,
                         "issue_slots_frac": slots / peak if (slots and peak) else None,
                         "issue_slots_note": "each 64-bit VALU instruction (SQ_INSTS_VALU_INT64) counted twice",
                         "montgomery_products_model": products,
                         "montgomery_products_per_s": products / ksec}}


COUNT_LIMIT = 3 * 10**7


def count_pairs(start: int, count: int) -> int:
    """Candidate pairs the reference's goldbach_count loop visits: p = 2..n/2 for each n."""
    n = start + 2 * np.arange(count, dtype=np.int64)
    return int((n // 2 - 1).sum())


def counts_leg(ctx) -> dict:
    """Partition counts c(n): every even n in [4, 3e7], and 4096 consecutive n from 1e10."""
    ctx.count_partitions(4, 10**4)                       # warm
    cnt = (COUNT_LIMIT - 4) // 2 + 1
    device_sync()
    t0 = time.perf_counter()
    c = ctx.count_partitions(4, cnt)
    dt = time.perf_counter() - t0
    assert int(c[(10**7 - 4) // 2]) == 38807 and int(c.min()) >= 1     # c(1e7), OEIS A065577
    t1 = time.perf_counter()
    hi = ctx.count_partitions(10**10, 4096)
    dt_hi = time.perf_counter() - t1
    assert int(hi[0]) == 18200488                                      # c(1e10)
    pairs = count_pairs(4, cnt)
    return {"workload": "c(n) for every even n in [4, 3e7]", "n": cnt, "seconds": dt,
            "value": pairs / dt, "unit": "pairs/s", "n_per_s": cnt / dt,
            "near_1e10": {"n": 4096, "seconds": dt_hi, "pairs_per_s": count_pairs(10**10, 4096) / dt_hi}}


def cpu_counts_baseline() -> dict:
    """The oracle's c(n) (reference loop: byte table, p ascending) on 512 n just below 3e7."""
    from _helpers import Oracle

    orc = Oracle()
    threads = min(16, os.cpu_count() or 1)
    start = COUNT_LIMIT - 2 * 511
    t0 = time.perf_counter()
    orc.count_partitions(start, 512, threads)
    dt = time.perf_counter() - t0
    return {"value": count_pairs(start, 512) / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"c(n) for the 512 even n below 3e7 (sieve included), {threads} threads, {dt:.2f} s wall"}


BIG_CASES = {1000: 26981, 10000: 47717}               # RESULTS.md:145-148 (the 10^1000 row run on one thread)
BIG_REF_SECONDS = {1000: 0.363, 10000: 182.0}         # RESULTS.md:146,148: 20 threads, GMP


def big_leg(ctx) -> dict:
    """big_check on the device: the smallest p with 10^k - p a strong probable prime to 25 bases."""
    if hasattr(sys, "set_int_max_str_digits"):
        sys.set_int_max_str_digits(0)
    ctx.big_min_partition(10**100)                       # warm (code, first launches)
    out = {}
    for k, want in BIG_CASES.items():
        t0 = time.perf_counter()
        p, st = ctx.big_min_partition(10**k)
        dt = time.perf_counter() - t0
        assert p == want, (k, p)
        out[f"10^{k}"] = {"p": p, "seconds": dt, "kernel_ms": st["kernel_ms"], "strong_tests": st["mr_tasks"],
                          "words": st["words"], "reference_seconds_20_threads": BIG_REF_SECONDS[k],
                          "vs_reference": BIG_REF_SECONDS[k] / dt}
    return out


def cpu_big_baseline() -> dict:
    """GMP's mpz_probab_prime_p(10^1000 - p, 25), the reference's test (big_check.cpp:107), over the
    primes p <= 26,981 in order, on a thread pool (ctypes releases the GIL during the call)."""
    from concurrent.futures import ThreadPoolExecutor

    from _helpers import Gmp, load_bigcheck

    gmp = Gmp()
    primes = [p for p in load_bigcheck().primes_upto(BIG_CASES[1000])]
    n = 10**1000
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        verdicts = list(ex.map(lambda p: gmp.probab_prime(n - p, 25), primes))
    dt = time.perf_counter() - t0
    assert verdicts[-1] > 0 and not any(verdicts[:-1])
    return {"value": dt, "unit": "s for 10^1000", "cores": threads, "kind": "port",
            "sample": f"GMP {gmp.version} mpz_probab_prime_p(10^1000 - p, 25) for the {len(primes)} primes "
                      f"p <= 26,981, {threads} threads"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-windows", action="store_true", help="skip the high-window (Baillie-PSW) leg")
    ap.add_argument("--no-counts", action="store_true", help="skip the partition-count leg")
    ap.add_argument("--no-big", action="store_true", help="skip the big_check (10^1000, 10^10000) leg")
    ap.add_argument("--no-ref-cpu", action="store_true",
                    help="skip the reference's own cpu_goldbach 1000000000 (one thread, 10-60 s)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device-to-device (counter calibration for the traffic figure)")
    ap.add_argument("--lds-probe", action="store_true",
                    help="only run the LDS issue probes (for rocprofv3 SQ_INSTS_LDS / SQ_BUSY_CYCLES)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the VALU issue probes (for rocprofv3 SQ_INSTS_VALU / SQ_BUSY_CYCLES)")
    ap.add_argument("--profile-commands", metavar="DIR",
                    help="print the rocprofv3 runs behind profiles/*.json, output under DIR, and exit")
    ap.add_argument("--summarize-profiles", metavar="DIR",
                    help="rebuild profiles/*.json from the CSV output of those runs under DIR, and exit")
    args = ap.parse_args()
    if args.profile_commands:
        for name in PROFILE_RUNS:
            print(profile_command(name, args.profile_commands))
        return 0
    if args.summarize_profiles:
        return summarize_profiles(args.summarize_profiles)
    if args.lds_probe:
        print(json.dumps({"lds_probe": lds_probe(), "build_hash": build_hash()}))
        return 0
    if args.valu_probe:
        print(json.dumps({"valu_probe": valu_probe(), "build_hash": build_hash()}))
        return 0
    if args.copy_probe:
        import torch

        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build verifies on one GPU; run with --gpus 1 (multi-device sharding "
                                   "is a design note, not built)"}))
        return 2

    from _helpers import load_pkg

    gb = load_pkg()
    ctx = gb.Context(0, max_seg_evens=PLAIN_SEG)
    evens = (LIMIT - 4) // 2 + 1

    def batched():
        st = ctx.verify_range(4, LIMIT, P_SMALL, form=gb.GB_FORM_BATCHED)
        assert st["evens"] == evens and st["failures"] == 0 and st["phase2"] == 0
        return st

    dt, outs, per_step = timed_passes(batched, args.steps, args.warmup)
    value = evens * args.steps / dt
    kms = float(np.mean([o["kernel_ms"] for o in outs]))           # device time of one pass's launches
    launches = len(ctx.range_plan(4, LIMIT, P_SMALL))
    achieved_gbs = ALG_BYTES_PER_EVEN * evens / (kms * 1e-3) / 1e9

    probe = valu_probe()
    peak_valu = probe["v_add_u32_independent"]["wave_insts_per_s"]
    lprobe = lds_probe()
    peak_lds = lprobe["ds_and_b32_scattered"]["wave_insts_per_s"]
    prof = load_profile(TRAFFIC_FILE)         # counters of the current sources, or None
    traffic_pass = prof.get("hbm_bytes_per_pass") if prof else None
    traffic = traffic_pass / launches if traffic_pass else None           # HBM bytes per launch
    hbm = {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
           "alg_bytes_per_pass": ALG_BYTES_PER_EVEN * evens, "alg_bytes_per_launch": ALG_BYTES_PER_EVEN * evens / launches,
           "traffic_bytes_per_pass": traffic_pass, "traffic_bytes_per_launch": traffic}
    if prof and prof.get("sq_insts_valu_per_pass"):
        # the binding resource: VALU issue (counted wave-instructions over the live time)
        achieved_valu = prof["sq_insts_valu_per_pass"] / (kms * 1e-3)
        roofline = {"bound": "valu", "achieved": achieved_valu, "peak": peak_valu, "unit": "wave-instructions/s",
                    "frac": achieved_valu / peak_valu, "traffic": traffic, "hbm": hbm}
        if prof.get("sq_insts_lds_per_pass"):
            # the other half-used unit: wave-level LDS instructions (the sieve's marks and the
            # marking's reads) against the rate the LDS probe reaches with scattered ds_and_b32
            achieved_lds = prof["sq_insts_lds_per_pass"] / (kms * 1e-3)
            roofline["lds"] = {"achieved": achieved_lds, "peak": peak_lds, "unit": "wave-instructions/s",
                               "frac": achieved_lds / peak_lds,
                               "counted": "SQ_INSTS_LDS of the pass's launches, " + os.path.relpath(TRAFFIC_FILE, ROOT),
                               "peak_note": "scattered ds_and_b32 rate of gb_probe_lds, measured in this run"}
    else:
        roofline = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                    "valu": None, "valu_note": "no counter profile of the current sources"}
    roofline.update({"kernel": "wheel_kernel<512, 294912>", "launches_per_pass": launches, "kernel_ms_per_pass": kms,
                     "kernel_ms_per_launch": kms / launches, "build_hash": build_hash(),
                     "profile": os.path.relpath(TRAFFIC_FILE, ROOT)})

    plain = None
    if not args.no_plain:
        def plain_pass():
            st = ctx.verify_range(4, LIMIT, P_SMALL, form=gb.GB_FORM_PLAIN, seg_evens=PLAIN_SEG)
            assert st["evens"] == evens and st["failures"] == 0
            return st
        psteps = max(1, min(args.steps, 2))
        pdt, pouts, _ = timed_passes(plain_pass, psteps, 0)
        pk = float(np.mean([o["kernel_ms"] for o in pouts]))
        plain = {"value": evens * psteps / pdt, "ms_per_step": pdt * 1e3 / psteps,
                 "kernel_ms": pk, "host_overhead": pdt * 1e3 / psteps / pk - 1.0, "seg_evens": PLAIN_SEG,
                 "limit_1e11": plain_1e11(ctx, gb), "limit_1e9": limit_1e9(ctx, gb, not args.no_ref_cpu)}

    serve = None if args.no_serve else serving(ctx)
    high = None if args.no_windows else windows_leg(ctx, probe)
    counts = None if args.no_counts else counts_leg(ctx)
    big = None if args.no_big else big_leg(ctx)
    ctx.close()
    cpu = None if args.no_cpu else cpu_baseline()
    if counts is not None and not args.no_cpu:
        counts["cpu_baseline"] = cpu_counts_baseline()
    if big is not None and not args.no_cpu:
        big["cpu_baseline"] = cpu_big_baseline()
    if high is not None and not args.no_cpu:
        high["cpu_baseline"] = cpu_windows_baseline()
    if serve is not None and not args.no_cpu:
        serve["cpu_baseline"] = cpu_serving_baseline()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "evens/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REF_PUBLISHED_EVENS_PER_S,
        "dtype": "u64",
        "data": "synthetic (the input is the integer range itself)",
        "config": {"workload": "all even n in [4, 1e12], batched persistent form", "limit": LIMIT,
                   "p_small": P_SMALL, "tile_evens": 27 << 15, "workgroups_per_cu": 2,
                   "parallelism": "1 GPU",
                   "plain_seg_evens": PLAIN_SEG},
        "roofline": roofline,
        "step_spread": spread(per_step),
        "valu_probe": probe,
        "lds_probe": lprobe,
        "cpu_baseline": cpu,
        "plain": plain,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "serving": serve,
        "high_windows": high,
        "partition_counts": counts,
        "big_check": big,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
