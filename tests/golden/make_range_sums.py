#!/usr/bin/env python3
"""Writes range_sums.json: the statistics of whole ranges from the oracle (oracle/oracle.c,
orc_verify_range: per-n ascending prime scan, cpu_goldbach.cpp:19-22 order; Phase 1 with
P_SMALL = 1e6, Phase 2 to 1e8, goldbach.cu:316-373), for the device's range tests.

    python3 tests/golden/make_range_sums.py     # ranges already in the file are kept

The small ranges take ~20 s on 8 threads; [4, 1e12] takes about half an hour.

Each range's largest p_min is checked against the published records (T. Oliveira e Silva's
table of minimal Goldbach partitions, OEIS A025018/A025019) in tests/test_range_sums.py,
which pins the scan independently of the oracle."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _helpers import Oracle  # noqa: E402

RANGES = [(4, 10**9), (4, 10**10), (10**10, 10**10 + 2**31),
          (10**12 - 2**31, 10**12), (4, 10**12)]


def main() -> None:
    path = os.path.join(HERE, "range_sums.json")
    have = {}
    if os.path.exists(path):
        with open(path) as f:
            have = {(r["start"], r["limit"]): r for r in json.load(f)}
    orc = Oracle()
    for start, limit in RANGES:
        if (start, limit) in have:
            continue
        t = time.time()
        st = orc.verify_range(start, limit, 1 << 27, 10**6)
        have[(start, limit)] = {"start": start, "limit": limit, "p_small": 10**6, **st}
        print(have[(start, limit)], f"{time.time() - t:.1f} s", flush=True)
        with open(path, "w") as f:   # after every range, so a long run keeps what it finished
            json.dump([have[r] for r in RANGES if r in have], f, indent=1)


if __name__ == "__main__":
    main()
