"""Partition counts c(n) on the device (gb_count_partitions) against the oracle
and the reference's own goldbach_count (cpu_goldbach.cpp:29-36): ranges from
n = 4 (2 + 2, 3 + 3), ranges crossing the narrow kernel's blocks of 8192 n and
its LDS chunks of 2048 words (n near 2^20 * k), ranges of 65536 n and more (the
kernel whose workgroups take n 128 apart), ragged ends, published values
c(10^k) and argument errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C_POW10 = (2, 6, 28, 127, 810, 5402, 38807, 291400, 2274205, 18200488)   # OEIS A065577, k = 1..10


@pytest.mark.parametrize("start,count", [(4, 1), (4, 2), (6, 1), (4, 5000), (8, 4097), (1_000_000, 3000),
                                         (1_048_000, 700), (2_097_100, 130), (16_000_000, 64),
                                         # 65536 n and more: workgroups of n 128 apart (one shift of R each)
                                         (4, 65536), (6, 65536), (4, 70001), (123_456, 131_073),
                                         (2_000_002, 65600), (3_000_000, 66000)])
def test_ranges_vs_oracle(ctx, orc, start, count):
    got = ctx.count_partitions(start, count)
    ref = orc.count_partitions(start, count)
    assert np.array_equal(got, ref), (start, count, np.nonzero(got != ref)[0][:5])


def test_block_and_chunk_edges(ctx, orc):
    # blocks of 8192 n: ranges ending one below, at and one above a block edge
    for start in (4, 8190, 8192, 8194):
        for count in (4095, 4096, 4097, 8193):
            assert np.array_equal(ctx.count_partitions(start, count), orc.count_partitions(start, count)), \
                (start, count)


def test_published_values(ctx):
    for k, want in enumerate(C_POW10, start=1):
        assert int(ctx.count_partitions(10**k, 1)[0]) == want, k


def test_matches_reference_goldbach_count(ctx, ref):
    for start, count in ((4, 2000), (500_002, 300)):
        assert np.array_equal(ctx.count_partitions(start, count), ref.goldbach_count(start, count)), start


def test_argument_errors(ctx, gb):
    for start, count in ((3, 1), (5, 4), (2, 1), (0, 3), (2**36 + 2, 1), (2**36 - 4, 4)):
        with pytest.raises(gb.GoldbachError):
            ctx.count_partitions(start, count)
    assert ctx.count_partitions(4, 0).size == 0
