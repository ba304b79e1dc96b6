#!/usr/bin/env python3
"""Writes range_sums.json: the statistics of whole ranges from the oracle (oracle/oracle.c,
orc_verify_range: per-n ascending prime scan, cpu_goldbach.cpp:19-22 order; Phase 1 with
P_SMALL = 1e6, Phase 2 to 1e8, goldbach.cu:316-373), for the device's range tests.

    python3 tests/golden/make_range_sums.py     # ~20 s on 8 threads

[4, 1e10]'s largest p_min, 1877 at n = 7,473,202,036, is a published record (T. Oliveira e
Silva's table of minimal Goldbach partitions, OEIS A025018/A025019), which pins the scan
independently of the oracle."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _helpers import Oracle  # noqa: E402

RANGES = [(4, 10**9), (4, 10**10), (10**10, 10**10 + 2**31)]


def main() -> None:
    orc = Oracle()
    out = []
    for start, limit in RANGES:
        st = orc.verify_range(start, limit, 1 << 27, 10**6)
        out.append({"start": start, "limit": limit, "p_small": 10**6, **st})
        print(out[-1], flush=True)
    with open(os.path.join(HERE, "range_sums.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
