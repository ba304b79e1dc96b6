"""Python view of the gfx950 Goldbach verifier's C ABI (include/goldbach_gpu.h).

The product is the HIP library ``lib/libgoldbach_gpu.so`` built from ``csrc/``;
this module only binds it with ctypes so tests, the benchmark and Python
callers can drive it.  Nothing here computes primes: when the library is
missing, loading fails loudly instead of falling back to any CPU path.

Entry points mirror the reference's per-segment loop (src/goldbach.cu:420-485),
its sieve (src/goldbach.cu:198-259) and single_check (src/single_check.cu).
"""
from __future__ import annotations

import ctypes
import os
import time
import weakref
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libgoldbach_gpu.so")
HOST_LIB_PATH = os.path.join(PKG_DIR, "lib", "libgoldbach.so")
BIN_DIR = os.path.join(PKG_DIR, "bin")

GB_OK, GB_EINVAL, GB_ENOMEM, GB_EDEVICE, GB_ENODEV, GB_ERANGE, GB_EAGAIN, GB_EIO = range(8)
GB_WANT_PMIN = 0x1
GB_TUNE_AUTO = 0x2
GB_FORM_PLAIN = 0
GB_FORM_BATCHED = 1

# Every symbol include/goldbach_gpu.h declares.
EXPORTS = (
    "gb_open", "gb_close", "gb_last_error", "gb_device_count", "gb_device_info",
    "gb_sieve_window", "gb_verify_segment", "gb_verify_range", "gb_verify_windows",
    "gb_min_partition_batch", "gb_autotune", "gb_count_partitions", "gb_verify_range_logged", "gb_log_read",
    "gb_big_min_partition", "gb_big_from_decimal", "gb_big_to_decimal",
    "gb_serve_open", "gb_serve_submit", "gb_serve_poll", "gb_serve_stats_get", "gb_serve_close",
    "gb_serve_last_error", "gb_range_plan", "gb_verify_segment_noalloc", "gb_kernel_times_get",
)
GB_CFG_KERNEL_TIMES = 0x1
GB_KERNEL_PLAIN, GB_KERNEL_WHEEL, GB_KERNEL_FUSED = 0, 1, 2

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32


class GbConfig(ctypes.Structure):
    _fields_ = [("max_n", u64), ("max_seg_evens", u64), ("phase2_limit", u64), ("mem_budget", u64),
                ("serve_slots", u32), ("flags", u32)]


class GbKernelTimes(ctypes.Structure):
    _fields_ = [("segments", u64), ("window_bits", u64), ("sieve_primes", u64), ("evens", u64),
                ("sieve_ms", ctypes.c_double), ("mark_ms", ctypes.c_double), ("other_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GbSegStats(ctypes.Structure):
    _fields_ = [("evens", u64), ("phase2", u64), ("failures", u64), ("first_fail_n", u64),
                ("sum_pmin", u64), ("max_pmin", u64), ("argmax_n", u64), ("kernel_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GbTuning(ctypes.Structure):
    _fields_ = [("form", u32), ("tile_log2", u32), ("seg_evens", u64), ("wg_per_cu", u32), ("flags", u32),
                ("lane_item_div", u32), ("wave_item_div", u32), ("mr_proof_stride", u32)]


class GbChunk(ctypes.Structure):
    _fields_ = [("first_n", u64), ("last_n", u64), ("tiles", u64), ("kernel", u32), ("tile_log2", u32),
                ("wg_per_cu", u32), ("reserved", u32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class GbWindow(ctypes.Structure):
    _fields_ = [("start", u64), ("evens", u64)]


class GbServeConfig(ctypes.Structure):
    _fields_ = [("slots", u32), ("chunks_per_step", u32), ("mem_budget", u64), ("queue_cap", u64)]


class GbServeStats(ctypes.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("steps", u64), ("step_queries", u64),
                ("max_resident", u64), ("queued", u64), ("resident", u64), ("slots", u32), ("reserved", u32),
                ("queue_ns", u64), ("admit_ns", u64), ("device_ns", u64), ("retire_ns", u64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class GbLogHeader(ctypes.Structure):
    _fields_ = [("start", u64), ("limit", u64), ("p_small", u64), ("phase2_limit", u64), ("chunk_evens", u64),
                ("want_pmin", u32), ("reserved", u32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class GbLogRecord(ctypes.Structure):
    _fields_ = [(name, u64) for name in ("chunk", "first_n", "last_n", "evens", "phase2", "failures",
                                          "first_fail_n", "sum_pmin", "max_pmin", "argmax_n")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GbLogInfo(ctypes.Structure):
    _fields_ = [("resumed_chunks", u64), ("resumed_evens", u64), ("done_chunks", u64), ("total_chunks", u64),
                ("complete", u32), ("truncated", u32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GbBigStats(ctypes.Structure):
    _fields_ = [("primes_scanned", u64), ("candidates", u64), ("mr_tasks", u64), ("batches", u32), ("words", u32),
                ("kernel_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# gb_result as a numpy record (48 bytes, same field order and widths)
RESULT_DTYPE = np.dtype([("id", "<u8"), ("n", "<u8"), ("p", "<u8"), ("submit_ns", "<u8"), ("done_ns", "<u8"),
                         ("steps", "<u4"), ("reserved", "<u4")])


class GoldbachError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.message = message


_lib = None


def library() -> ctypes.CDLL:
    """Load lib/libgoldbach_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "gb_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(GbConfig), ctypes.POINTER(vp)]),
        "gb_close": (ctypes.c_int, [vp]),
        "gb_last_error": (ctypes.c_char_p, [vp]),
        "gb_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
        "gb_device_info": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(u64),
                                          ctypes.POINTER(ctypes.c_int)]),
        "gb_sieve_window": (ctypes.c_int, [vp, u64, u64, vp]),
        "gb_verify_segment": (ctypes.c_int, [vp, u64, u64, u64, u32, vp, vp, ctypes.POINTER(GbSegStats)]),
        "gb_verify_segment_noalloc": (ctypes.c_int, [vp, u64, u64, u64, u32, vp, vp, ctypes.POINTER(GbSegStats)]),
        "gb_verify_range": (ctypes.c_int, [vp, u64, u64, u64, ctypes.POINTER(GbTuning), ctypes.POINTER(GbSegStats),
                                           vp]),
        "gb_verify_windows": (ctypes.c_int, [vp, ctypes.POINTER(GbWindow), u32, u64, u32,
                                             ctypes.POINTER(GbSegStats), ctypes.POINTER(GbSegStats)]),
        "gb_min_partition_batch": (ctypes.c_int, [vp, vp, u32, vp]),
        "gb_count_partitions": (ctypes.c_int, [vp, u64, u64, vp]),
        "gb_verify_range_logged": (ctypes.c_int, [vp, u64, u64, u64, ctypes.POINTER(GbTuning), ctypes.c_char_p, u64,
                                                  u64, ctypes.POINTER(GbSegStats), ctypes.POINTER(GbLogInfo), vp]),
        "gb_big_min_partition": (ctypes.c_int, [vp, vp, u32, u32, u64, ctypes.POINTER(u64),
                                                ctypes.POINTER(GbBigStats)]),
        "gb_big_from_decimal": (ctypes.c_int, [ctypes.c_char_p, vp, u32, ctypes.POINTER(u32)]),
        "gb_big_to_decimal": (ctypes.c_int, [vp, u32, ctypes.c_char_p, ctypes.c_size_t]),
        "gb_log_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(GbLogHeader), ctypes.POINTER(GbLogRecord),
                                       u64, ctypes.POINTER(u64), ctypes.POINTER(u32)]),
        "gb_autotune": (ctypes.c_int, [vp, u64, u64, u64, u32, u32, ctypes.POINTER(GbTuning),
                                       ctypes.POINTER(ctypes.c_double)]),
        "gb_kernel_times_get": (ctypes.c_int, [vp, ctypes.POINTER(GbKernelTimes), ctypes.c_int]),
        "gb_range_plan": (ctypes.c_int, [vp, u64, u64, u64, ctypes.POINTER(GbTuning), ctypes.POINTER(GbChunk), u64,
                                         ctypes.POINTER(u64)]),
        "gb_serve_open": (ctypes.c_int, [vp, ctypes.POINTER(GbServeConfig), ctypes.POINTER(vp)]),
        "gb_serve_submit": (ctypes.c_int, [vp, vp, u32, ctypes.POINTER(u64)]),
        "gb_serve_poll": (ctypes.c_int, [vp, vp, u32, ctypes.POINTER(u32), u32]),
        "gb_serve_stats_get": (ctypes.c_int, [vp, ctypes.POINTER(GbServeStats)]),
        "gb_serve_close": (ctypes.c_int, [vp]),
        "gb_serve_last_error": (ctypes.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def big_words(n: int) -> np.ndarray:
    """n >= 0 as little-endian 32-bit words (at least one)."""
    if n < 0:
        raise ValueError("n must be >= 0")
    nw = max(1, (n.bit_length() + 31) // 32)
    return np.frombuffer(n.to_bytes(4 * nw, "little"), dtype=np.uint32).copy()


def from_decimal(digits: str, cap: int = 1 << 16) -> np.ndarray:
    """gb_big_from_decimal: decimal digits to little-endian 32-bit words."""
    w = np.zeros(cap, dtype=np.uint32)
    nw = u32(0)
    rc = library().gb_big_from_decimal(digits.encode(), _ptr(w), cap, ctypes.byref(nw))
    if rc != GB_OK:
        raise GoldbachError(rc, f"gb_big_from_decimal failed for {digits[:20]!r}")
    return w[:nw.value].copy()


def to_decimal(words) -> str:
    """gb_big_to_decimal: little-endian 32-bit words to decimal digits."""
    w = np.ascontiguousarray(np.asarray(words, dtype=np.uint32))
    buf = ctypes.create_string_buffer(10 * len(w) + 2)
    rc = library().gb_big_to_decimal(_ptr(w), len(w), buf, len(buf))
    if rc != GB_OK:
        raise GoldbachError(rc, "gb_big_to_decimal failed")
    return buf.value.decode()


def read_log(path: str) -> tuple:
    """(header dict, list of record dicts, torn) of a run log; needs no device."""
    lib = library()
    hdr = GbLogHeader()
    n = u64(0)
    torn = u32(0)
    rc = lib.gb_log_read(os.fsencode(path), ctypes.byref(hdr), None, 0, ctypes.byref(n), ctypes.byref(torn))
    if rc != GB_OK:
        raise GoldbachError(rc, f"cannot read run log {path}")
    recs = (GbLogRecord * max(n.value, 1))()
    rc = lib.gb_log_read(os.fsencode(path), ctypes.byref(hdr), recs, n.value, ctypes.byref(n), ctypes.byref(torn))
    if rc != GB_OK:
        raise GoldbachError(rc, f"cannot read run log {path}")
    return hdr.as_dict(), [recs[i].as_dict() for i in range(n.value)], bool(torn.value)


def device_count() -> int:
    n = ctypes.c_int(0)
    library().gb_device_count(ctypes.byref(n))
    return n.value


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class SegmentResult:
    stats: dict
    verified: Optional[np.ndarray] = None   # uint64 words, bit j of word w <-> n = start + 2(64w+j)
    pmin: Optional[np.ndarray] = None       # uint32 per even n


class Context:
    """One device context: resident prime tables, stream and segment buffers."""

    def __init__(self, device: int = 0, max_seg_evens: int = 0, phase2_limit: int = 0, max_n: int = 0,
                 mem_budget: int = 0, kernel_times: bool = False):
        self._lib = library()
        cfg = GbConfig(max_n, max_seg_evens, phase2_limit, mem_budget, 0,
                       GB_CFG_KERNEL_TIMES if kernel_times else 0)
        h = ctypes.c_void_p()
        rc = self._lib.gb_open(device, ctypes.byref(cfg), ctypes.byref(h))
        if rc != GB_OK:
            raise GoldbachError(rc, self._lib.gb_last_error(None).decode())
        self._h = h
        self._servers = weakref.WeakSet()

    # -- lifecycle -------------------------------------------------------
    def close(self) -> None:
        for srv in list(getattr(self, "_servers", ())):
            srv.close()
        if getattr(self, "_h", None):
            self._lib.gb_close(self._h)
            self._h = None

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != GB_OK:
            raise GoldbachError(rc, self._lib.gb_last_error(self._h).decode())

    def kernel_times(self, reset: bool = True) -> dict:
        """Per-kernel device time of the plain form since the last reset (kernel_times=True contexts)."""
        kt = GbKernelTimes()
        self._check(self._lib.gb_kernel_times_get(self._h, ctypes.byref(kt), 1 if reset else 0))
        return kt.as_dict()

    # -- sieve -----------------------------------------------------------
    def sieve_window(self, q_low: int, q_high: int) -> np.ndarray:
        nbits = (q_high - q_low) // 2 + 1 if q_high >= q_low else 0
        out = np.zeros(max((nbits + 63) // 64, 1), dtype=np.uint64)
        self._check(self._lib.gb_sieve_window(self._h, q_low, q_high, _ptr(out)))
        return out

    # -- plain per-segment form -------------------------------------------
    def verify_segment(self, seg_start: int, count: int, p_small: int, want_pmin: bool = False,
                       want_verified: bool = True, want_pmin_array: bool = False) -> SegmentResult:
        st = GbSegStats()
        ver = np.zeros((count + 63) // 64, dtype=np.uint64) if want_verified else None
        pm = np.zeros(count, dtype=np.uint32) if want_pmin_array else None
        flags = GB_WANT_PMIN if (want_pmin or want_pmin_array) else 0
        self._check(self._lib.gb_verify_segment(self._h, seg_start, count, p_small, flags,
                                                _ptr(ver) if ver is not None else None,
                                                _ptr(pm) if pm is not None else None, ctypes.byref(st)))
        return SegmentResult(st.as_dict(), ver, pm)

    def verify_segment_noalloc(self, seg_start: int, count: int, p_small: int, want_pmin: bool = False,
                               want_verified: bool = True, want_pmin_array: bool = False) -> SegmentResult:
        """verify_segment through gb_verify_segment_noalloc (static device scratch, count <= 2^20)."""
        st = GbSegStats()
        ver = np.zeros((count + 63) // 64, dtype=np.uint64) if want_verified else None
        pm = np.zeros(count, dtype=np.uint32) if want_pmin_array else None
        flags = GB_WANT_PMIN if (want_pmin or want_pmin_array) else 0
        self._check(self._lib.gb_verify_segment_noalloc(self._h, seg_start, count, p_small, flags,
                                                        _ptr(ver) if ver is not None else None,
                                                        _ptr(pm) if pm is not None else None, ctypes.byref(st)))
        return SegmentResult(st.as_dict(), ver, pm)

    # -- whole range -------------------------------------------------------
    def verify_range(self, start: int, limit: int, p_small: int, form: int = GB_FORM_BATCHED,
                     tile_log2: int = 0, seg_evens: int = 0, wg_per_cu: int = 0, want_pmin: bool = False,
                     autotune: bool = False, progress: Optional[ctypes.c_uint64] = None,
                     lane_item_div: int = 0, wave_item_div: int = 0, mr_proof_stride: int = 0) -> dict:
        """progress (optional): a ctypes.c_uint64 the library updates with the evens finished."""
        flags = (GB_WANT_PMIN if want_pmin else 0) | (GB_TUNE_AUTO if autotune else 0)
        tun = GbTuning(form, tile_log2, seg_evens, wg_per_cu, flags, lane_item_div, wave_item_div, mr_proof_stride)
        st = GbSegStats()
        self._check(self._lib.gb_verify_range(self._h, start, limit, p_small, ctypes.byref(tun), ctypes.byref(st),
                                              ctypes.byref(progress) if progress is not None else None))
        return st.as_dict()

    def range_plan(self, start: int, limit: int, p_small: int = 1_000_000, form: int = GB_FORM_BATCHED,
                   tile_log2: int = 0, seg_evens: int = 0, wg_per_cu: int = 0, autotune: bool = False) -> list:
        """gb_range_plan: the segments or launches verify_range would run, as dicts."""
        tun = GbTuning(form, tile_log2, seg_evens, wg_per_cu, GB_TUNE_AUTO if autotune else 0)
        n = u64(0)
        self._check(self._lib.gb_range_plan(self._h, start, limit, p_small, ctypes.byref(tun), None, 0,
                                            ctypes.byref(n)))
        arr = (GbChunk * max(n.value, 1))()
        self._check(self._lib.gb_range_plan(self._h, start, limit, p_small, ctypes.byref(tun), arr, n.value,
                                            ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def verify_range_logged(self, start: int, limit: int, p_small: int, log_path: str, chunk_evens: int = 0,
                            max_chunks: int = 0, form: int = GB_FORM_BATCHED, tile_log2: int = 0,
                            seg_evens: int = 0, wg_per_cu: int = 0, want_pmin: bool = False) -> tuple:
        """(stats over every chunk in the log, log info): resumes from log_path where it stopped."""
        tun = GbTuning(form, tile_log2, seg_evens, wg_per_cu, GB_WANT_PMIN if want_pmin else 0)
        st = GbSegStats()
        info = GbLogInfo()
        self._check(self._lib.gb_verify_range_logged(self._h, start, limit, p_small, ctypes.byref(tun),
                                                     os.fsencode(log_path), chunk_evens, max_chunks,
                                                     ctypes.byref(st), ctypes.byref(info), None))
        return st.as_dict(), info.as_dict()

    def big_min_partition(self, n: int, rounds: int = 25, p_max: int = 0) -> tuple:
        """(smallest prime p <= p_max with n - p a strong probable prime to the first `rounds`
        prime bases, 0 if none; stats dict) for an even n >= 4 of any size up to 73,728 bits."""
        w = big_words(n)
        p = u64(0)
        st = GbBigStats()
        self._check(self._lib.gb_big_min_partition(self._h, _ptr(w), len(w), rounds, p_max, ctypes.byref(p),
                                                   ctypes.byref(st)))
        return p.value, st.as_dict()

    def autotune(self, start: int, limit: int, p_small: int, form: int = GB_FORM_BATCHED,
                 want_pmin: bool = False) -> tuple:
        """(tuning dict, evens/s on the samples; 0 when cached or the range is too short to sample)."""
        t = GbTuning()
        rate = ctypes.c_double(0)
        self._check(self._lib.gb_autotune(self._h, start, limit, p_small, form, GB_WANT_PMIN if want_pmin else 0,
                                          ctypes.byref(t), ctypes.byref(rate)))
        return ({"form": t.form, "tile_log2": t.tile_log2, "seg_evens": t.seg_evens, "wg_per_cu": t.wg_per_cu,
                 "flags": t.flags, "lane_item_div": t.lane_item_div, "wave_item_div": t.wave_item_div}, rate.value)

    def verify_windows(self, windows: Sequence[tuple], p_small: int, want_pmin: bool = False):
        arr = (GbWindow * len(windows))(*[GbWindow(int(a), int(e)) for a, e in windows])
        per = (GbSegStats * len(windows))()
        tot = GbSegStats()
        self._check(self._lib.gb_verify_windows(self._h, arr, len(windows), p_small,
                                                GB_WANT_PMIN if want_pmin else 0, per, ctypes.byref(tot)))
        return [p.as_dict() for p in per], tot.as_dict()

    # -- single numbers ----------------------------------------------------
    def min_partition(self, ns) -> np.ndarray:
        n = np.ascontiguousarray(np.asarray(ns, dtype=np.uint64))
        out = np.zeros(n.shape[0], dtype=np.uint64)
        self._check(self._lib.gb_min_partition_batch(self._h, _ptr(n), n.shape[0], _ptr(out)))
        return out

    def count_partitions(self, start: int, count: int) -> np.ndarray:
        """c(n) for the count even n from start: primes p <= n/2 with n - p prime."""
        out = np.zeros(count, dtype=np.uint64)
        self._check(self._lib.gb_count_partitions(self._h, start, count, _ptr(out)))
        return out

    def server(self, slots: int = 0, chunks_per_step: int = 0, mem_budget: int = 0, queue_cap: int = 0) -> "Server":
        return Server(self, slots, chunks_per_step, mem_budget, queue_cap)


class Server:
    """Continuous-batching serving engine on a context (gb_serve_*)."""

    def __init__(self, ctx: Context, slots: int = 0, chunks_per_step: int = 0, mem_budget: int = 0,
                 queue_cap: int = 0):
        self._lib = ctx._lib
        self._ctx = ctx
        cfg = GbServeConfig(slots, chunks_per_step, mem_budget, queue_cap)
        h = ctypes.c_void_p()
        ctx._check(self._lib.gb_serve_open(ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._buf = np.zeros(65536, dtype=RESULT_DTYPE)
        self._stash = np.zeros(0, dtype=RESULT_DTYPE)   # results run() saw for other submissions
        ctx._servers.add(self)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gb_serve_close(self._h)
            self._h = None
            self._ctx._servers.discard(self)

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != GB_OK:
            raise GoldbachError(rc, self._lib.gb_serve_last_error(self._h).decode())

    def submit(self, ns) -> int:
        """Queue even n >= 4; returns the id of the first (the rest follow in order)."""
        n = np.ascontiguousarray(np.asarray(ns, dtype=np.uint64).reshape(-1))
        first = u64(0)
        self._check(self._lib.gb_serve_submit(self._h, _ptr(n), n.shape[0], ctypes.byref(first)))
        return first.value

    def poll(self, max_results: int = 65536, wait_ms: int = 0) -> np.ndarray:
        """Finished results (RESULT_DTYPE records) in completion order."""
        if self._stash.shape[0]:
            out, self._stash = self._stash[:max_results], self._stash[max_results:]
            return out
        return self._poll(max_results, wait_ms)

    def _poll(self, max_results: int, wait_ms: int) -> np.ndarray:
        if max_results > self._buf.shape[0]:
            self._buf = np.zeros(max_results, dtype=RESULT_DTYPE)
        got = u32(0)
        self._check(self._lib.gb_serve_poll(self._h, _ptr(self._buf), max_results, ctypes.byref(got), wait_ms))
        return self._buf[:got.value].copy()

    def stats(self) -> dict:
        st = GbServeStats()
        self._check(self._lib.gb_serve_stats_get(self._h, ctypes.byref(st)))
        return st.as_dict()

    def run(self, ns, timeout_s: float = 60.0) -> np.ndarray:
        """Submit ns and wait for all of them; returns p in submission order."""
        n = np.asarray(ns, dtype=np.uint64).reshape(-1)
        first = self.submit(n)
        out = np.zeros(n.shape[0], dtype=np.uint64)
        left = n.shape[0]
        t_end = time.monotonic() + timeout_s
        while left:
            r = self._poll(65536, wait_ms=50)
            mine = (r["id"] >= first) & (r["id"] < first + n.shape[0])
            out[(r["id"][mine] - first).astype(np.int64)] = r["p"][mine]
            self._stash = np.concatenate([self._stash, r[~mine]])
            left -= int(mine.sum())
            if time.monotonic() > t_end:
                raise TimeoutError(f"serving: {left} queries unfinished after {timeout_s} s")
        return out


def unpack_bits(words: np.ndarray, nbits: int) -> np.ndarray:
    """uint64 words (LSB-first) -> uint8 array of nbits flags."""
    b = np.unpackbits(words.view(np.uint8), bitorder="little")
    return b[:nbits]
