"""Shared loaders for the test-suite: the package under test (through its C ABI)
and the CPU oracle (test infrastructure only)."""
from __future__ import annotations

import ctypes
import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "goldbach-gpu_amd")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref")

u64 = ctypes.c_uint64


def load_pkg():
    """Import goldbach-gpu_amd/ (its directory name is not an identifier)."""
    name = "goldbach_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


class OrcStats(ctypes.Structure):
    _fields_ = [(n, u64) for n in ("evens", "phase2", "failures", "first_fail_n", "sum_pmin", "max_pmin",
                                   "argmax_n")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Oracle:
    """ctypes view of oracle/liboracle.so."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: build oracle/ first")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.orc_isqrt.restype = u64; L.orc_isqrt.argtypes = [u64]
        L.orc_simple_sieve.restype = ctypes.c_size_t; L.orc_simple_sieve.argtypes = [u64, vp, ctypes.c_size_t]
        L.orc_segmented_sieve.restype = ctypes.c_int; L.orc_segmented_sieve.argtypes = [u64, u64, vp]
        L.orc_bitset_word_count.restype = u64; L.orc_bitset_word_count.argtypes = [u64]
        L.orc_prime_bitset.restype = None; L.orc_prime_bitset.argtypes = [u64, vp]
        L.orc_window_word_count.restype = u64; L.orc_window_word_count.argtypes = [u64, u64]
        L.orc_window_bits.restype = ctypes.c_int; L.orc_window_bits.argtypes = [u64, u64, vp]
        L.orc_is_prime.restype = ctypes.c_int; L.orc_is_prime.argtypes = [u64]
        L.orc_strong_first.restype = ctypes.c_int; L.orc_strong_first.argtypes = [u64, ctypes.c_int]
        L.orc_count_partitions.restype = ctypes.c_int
        L.orc_count_partitions.argtypes = [u64, u64, ctypes.c_void_p, ctypes.c_int]
        L.orc_first_strong_pseudoprime.restype = u64
        L.orc_first_strong_pseudoprime.argtypes = [ctypes.c_int, u64]
        L.orc_min_partition.restype = u64; L.orc_min_partition.argtypes = [u64, u64]
        L.orc_min_partition_batch.restype = None
        L.orc_min_partition_batch.argtypes = [vp, ctypes.c_size_t, u64, vp, ctypes.c_int]
        L.orc_verify_segment.restype = ctypes.c_int
        L.orc_verify_segment.argtypes = [u64, u64, u64, u64, vp, vp, ctypes.POINTER(OrcStats)]
        L.orc_verify_range.restype = ctypes.c_int
        L.orc_verify_range.argtypes = [u64, u64, u64, u64, u64, ctypes.c_int, ctypes.POINTER(OrcStats)]
        L.orc_cpu_goldbach.restype = ctypes.c_int
        L.orc_cpu_goldbach.argtypes = [u64, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)]
        self.L = L

    def simple_sieve(self, limit):
        n = self.L.orc_simple_sieve(limit, None, 0)
        out = np.zeros(max(n, 1), dtype=np.uint64)
        self.L.orc_simple_sieve(limit, out.ctypes.data, n)
        return out[:n]

    def segmented_sieve(self, low, high):
        out = np.zeros(high - low + 1, dtype=np.uint8)
        assert self.L.orc_segmented_sieve(low, high, out.ctypes.data) == 0
        return out

    def prime_bitset(self, limit):
        nw = self.L.orc_bitset_word_count(limit)
        out = np.zeros(max(nw, 1), dtype=np.uint64)
        self.L.orc_prime_bitset(limit, out.ctypes.data)
        return out[:nw]

    def window_bits(self, q_low, q_high):
        nw = self.L.orc_window_word_count(q_low, q_high)
        out = np.zeros(max(nw, 1), dtype=np.uint64)
        assert self.L.orc_window_bits(q_low, q_high, out.ctypes.data) == 0
        return out[:nw]

    def strong_first(self, n, k):
        return bool(self.L.orc_strong_first(n, k))

    def first_strong_pseudoprime(self, k, limit):
        return int(self.L.orc_first_strong_pseudoprime(k, limit))

    def count_partitions(self, start, count, threads=0):
        out = np.zeros(count, dtype=np.uint64)
        assert self.L.orc_count_partitions(start, count, out.ctypes.data, threads) == 0
        return out

    def is_prime(self, n):
        return bool(self.L.orc_is_prime(n))

    def min_partition(self, n, bound=10**8):
        return int(self.L.orc_min_partition(n, bound))

    def min_partition_batch(self, ns, bound=10**8, threads=0):
        ns = np.ascontiguousarray(np.asarray(ns, dtype=np.uint64))
        out = np.zeros(ns.shape[0], dtype=np.uint64)
        self.L.orc_min_partition_batch(ns.ctypes.data, ns.shape[0], bound, out.ctypes.data, threads)
        return out

    def verify_segment(self, A, count, p_small, phase2_limit=10**8, want_bits=True, want_pmin=True):
        st = OrcStats()
        bits = np.zeros((count + 63) // 64, dtype=np.uint64) if want_bits else None
        pm = np.zeros(count, dtype=np.uint32) if want_pmin else None
        rc = self.L.orc_verify_segment(A, count, p_small, phase2_limit,
                                       bits.ctypes.data if bits is not None else None,
                                       pm.ctypes.data if pm is not None else None, ctypes.byref(st))
        assert rc == 0
        return st.as_dict(), bits, pm

    def verify_range(self, start, limit, seg_evens, p_small, phase2_limit=10**8, threads=0):
        st = OrcStats()
        assert self.L.orc_verify_range(start, limit, seg_evens, p_small, phase2_limit, threads,
                                       ctypes.byref(st)) == 0
        return st.as_dict()

    def cpu_goldbach(self, limit):
        c, f, ff = u64(), u64(), u64()
        assert self.L.orc_cpu_goldbach(limit, ctypes.byref(c), ctypes.byref(f), ctypes.byref(ff)) == 0
        return c.value, f.value, ff.value


class Ref:
    """The reference's own CPU sources, compiled by oracle/Makefile into oracle/_ref."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.ref_simple_sieve.restype = ctypes.c_size_t; L.ref_simple_sieve.argtypes = [u64, vp, ctypes.c_size_t]
        L.ref_segmented_sieve.restype = ctypes.c_int; L.ref_segmented_sieve.argtypes = [u64, u64, vp]
        L.ref_prime_bitset.restype = u64; L.ref_prime_bitset.argtypes = [u64, vp, u64, ctypes.c_int]
        L.ref_goldbach_count.restype = ctypes.c_int; L.ref_goldbach_count.argtypes = [u64, u64, vp]
        self.L = L

    def simple_sieve(self, limit):
        n = self.L.ref_simple_sieve(limit, None, 0)
        out = np.zeros(max(n, 1), dtype=np.uint64)
        self.L.ref_simple_sieve(limit, out.ctypes.data, n)
        return out[:n]

    def segmented_sieve(self, low, high):
        out = np.zeros(high - low + 1, dtype=np.uint8)
        self.L.ref_segmented_sieve(low, high, out.ctypes.data)
        return out

    def prime_bitset(self, limit, threads=1):
        nw = self.L.ref_prime_bitset(limit, None, 0, threads)
        out = np.zeros(max(nw, 1), dtype=np.uint64)
        self.L.ref_prime_bitset(limit, out.ctypes.data, nw, threads)
        return out[:nw]

    def goldbach_count(self, start, count):
        out = np.zeros(count, dtype=np.uint64)
        self.L.ref_goldbach_count(start, count, out.ctypes.data)
        return out


def unpack(words, nbits):
    return np.unpackbits(np.ascontiguousarray(words).view(np.uint8), bitorder="little")[:nbits]


def load_bigcheck():
    """oracle/bigcheck.py: the big-n search restated in Python (test infrastructure)."""
    name = "oracle_bigcheck"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "oracle", "bigcheck.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


class Gmp:
    """mpz_probab_prime_p of the image's GMP (the reference's big_check dependency,
    big_check.cpp:107), through ctypes.  mpz_t is {int alloc; int size; limb *d}."""

    class _Mpz(ctypes.Structure):
        _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]

    def __init__(self):
        lib = ctypes.CDLL("libgmp.so.10")
        P = ctypes.POINTER(self._Mpz)
        fn = lambda name: getattr(lib, name)       # names start with two underscores
        self._init, self._clear = fn("__gmpz_init"), fn("__gmpz_clear")
        self._set_str, self._pp = fn("__gmpz_set_str"), fn("__gmpz_probab_prime_p")
        self._init.argtypes = [P]
        self._clear.argtypes = [P]
        self._set_str.argtypes = [P, ctypes.c_char_p, ctypes.c_int]
        self._set_str.restype = ctypes.c_int
        self._pp.argtypes = [P, ctypes.c_int]
        self._pp.restype = ctypes.c_int
        self.version = ctypes.c_char_p.in_dll(lib, "__gmp_version").value.decode()

    def probab_prime(self, q: int, reps: int = 25) -> int:
        z = self._Mpz()
        self._init(ctypes.byref(z))
        try:
            assert self._set_str(ctypes.byref(z), format(q, "x").encode(), 16) == 0
            return self._pp(ctypes.byref(z), reps)
        finally:
            self._clear(ctypes.byref(z))
