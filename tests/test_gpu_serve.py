"""Serving loop with continuous batching (gb_serve_*): each query's answer is
the minimal partition p, identical whether the query runs alone, mixed into a
stream of arrivals of every magnitude, or through gb_min_partition_batch, and
equal to the oracle's.  The reference's single_check (src/single_check.cu:
136-282) answers one n per process with the first p a GPU thread finds; the
minimal p is what the oracle restates (cpu_goldbach.cpp:19-22 order)."""
import time

import numpy as np
import pytest

from _helpers import load_pkg

pytestmark = pytest.mark.gpu

# p_min(3325581707333960528) = 9781: about 1200 candidates, i.e. many steps
LONG_N = 3325581707333960528


def serving_ns(count, seed=1):
    """Even n log-uniform in [1e6, 2^64-2] (the serving workload's draw)."""
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(1e6), np.log(2.0**64 - 2), size=count))
    n = (x / 2).astype(np.uint64) * np.uint64(2)
    return np.maximum(n, np.uint64(4))


def collect(srv, want, timeout_s=60.0):
    out = []
    got = 0
    t_end = time.monotonic() + timeout_s
    while got < want:
        r = srv.poll(65536, wait_ms=50)
        out.append(r)
        got += r.shape[0]
        assert time.monotonic() < t_end, f"{want - got} queries unfinished"
    return np.concatenate(out)


def test_results_alone_and_in_the_mix(ctx, orc):
    special = np.array([4, 6, 8, 10, 98, 10**12, 2**64 - 2, LONG_N, 10**19, 60119912, 1_000_002],
                       dtype=np.uint64)
    ns = np.concatenate([special, serving_ns(3000)])
    want = orc.min_partition_batch(ns)
    batch = ctx.min_partition(ns)
    assert np.array_equal(batch, want)
    with ctx.server(chunks_per_step=1) as srv:
        # alone: one query on the device at a time
        alone = np.array([srv.run([n])[0] for n in ns[:300]], dtype=np.uint64)
        assert np.array_equal(alone, want[:300])
        # mixed: groups of random size arrive while earlier ones are in flight
        rng = np.random.default_rng(7)
        ids = np.zeros(ns.shape[0], dtype=np.uint64)
        i = 0
        results = []
        while i < ns.shape[0]:
            k = int(rng.integers(1, 200))
            first = srv.submit(ns[i:i + k])
            ids[i:i + k] = first + np.arange(min(k, ns.shape[0] - i), dtype=np.uint64)
            i += k
            results.append(srv.poll(65536))
        got = np.concatenate(results)
        r = np.concatenate([got, collect(srv, ns.shape[0] - got.shape[0])])
        assert r.shape[0] == ns.shape[0]
        order = np.argsort(r["id"])
        r = r[order]
        assert np.array_equal(r["id"], ids)
        assert np.array_equal(r["n"], ns)
        assert np.array_equal(r["p"], want)
        assert np.all(r["done_ns"] >= r["submit_ns"])
        st = srv.stats()
        assert st["completed"] == st["submitted"] == 300 + ns.shape[0]
        assert st["max_resident"] > 1
    # both ends of the pair are prime (single_check's contract)
    for n, p in zip(ns[:40], want[:40]):
        assert orc.is_prime(int(p)) and orc.is_prime(int(n) - int(p))


def test_server_after_and_before_plain_ranges(gb, orc):
    # A plain range of several segments runs its second segment stream; opening a server on the
    # same context releases that stream (so the server's four lane streams keep hardware queues
    # of their own), and the next plain range creates it again.  Answers stay exact throughout.
    with gb.Context(0, max_seg_evens=1 << 20) as c:
        want_range = orc.verify_range(4, 10**7, 1 << 20, 10**6)
        keys = ("evens", "phase2", "failures", "sum_pmin", "max_pmin", "argmax_n")
        for _ in range(2):
            got = c.verify_range(4, 10**7, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=1 << 20, want_pmin=True)
            assert all(got[k] == want_range[k] for k in keys), (got, want_range)
            ns = serving_ns(500, seed=3)
            with c.server() as srv:
                assert np.array_equal(srv.run(ns), orc.min_partition_batch(ns))
        got = c.verify_range(4, 10**7, 10**6, form=gb.GB_FORM_PLAIN, seg_evens=1 << 20, want_pmin=True)
        assert all(got[k] == want_range[k] for k in keys)


def test_long_query_keeps_its_slot_while_others_pass(ctx, orc):
    with ctx.server(chunks_per_step=1) as srv:
        long_id = srv.submit([LONG_N])
        short = serving_ns(64, seed=3)
        srv.submit(short)
        r = collect(srv, 65)
        by_id = {int(x["id"]): x for x in r}
        assert int(by_id[long_id]["p"]) == 9781
        assert int(by_id[long_id]["steps"]) >= 10   # 9781 is the ~1200th odd prime: 19 chunks of 64
        # queries submitted after it finished before it: they shared steps with it
        done_long = int(by_id[long_id]["done_ns"])
        assert sum(1 for x in r if int(x["id"]) != long_id and int(x["done_ns"]) < done_long) > 0
        assert np.array_equal(np.array([by_id[long_id + 1 + j]["p"] for j in range(64)], dtype=np.uint64),
                              orc.min_partition_batch(short))


def test_slots_bound_residency(ctx, orc):
    ns = serving_ns(1000, seed=5)
    with ctx.server(slots=8) as srv:
        p = srv.run(ns)
        st = srv.stats()
    assert np.array_equal(p, orc.min_partition_batch(ns))
    assert st["slots"] == 8 and st["max_resident"] <= 8 and st["steps"] >= 1000 // 8


def test_memory_budget_admission(ctx, orc):
    gb = load_pkg()
    # 16 bytes of device memory per resident query: an 80-byte budget admits 5 at a time
    ns = serving_ns(50, seed=9)
    with ctx.server(slots=1000, mem_budget=80) as srv:
        assert srv.stats()["slots"] == 5
        assert np.array_equal(srv.run(ns), orc.min_partition_batch(ns))
        assert srv.stats()["max_resident"] <= 5
    with pytest.raises(gb.GoldbachError) as e:
        ctx.server(mem_budget=15)
    assert e.value.status == gb.GB_ENOMEM
    # the slot state comes out of the context's own device budget
    with gb.Context(0, mem_budget=256 << 20) as c:
        with pytest.raises(gb.GoldbachError) as e:
            c.server(slots=1 << 24)                      # 256 MiB of slot state
        assert e.value.status == gb.GB_ENOMEM and "budget" in e.value.message
        with c.server(slots=1 << 16) as srv:
            assert list(srv.run([98, 100])) == [19, 3]


def test_submit_contract(ctx):
    gb = load_pkg()
    with ctx.server(queue_cap=10, slots=1) as srv:
        with pytest.raises(gb.GoldbachError) as e:
            srv.submit([10, 7, 12])
        assert e.value.status == gb.GB_EINVAL
        with pytest.raises(gb.GoldbachError) as e:
            srv.submit([2])
        assert e.value.status == gb.GB_EINVAL
        assert srv.stats()["submitted"] == 0
        with pytest.raises(gb.GoldbachError) as e:
            srv.submit([10] * 11)
        assert e.value.status == gb.GB_EAGAIN
        assert list(srv.run([10, 12])) == [3, 5]
        assert srv.poll().shape[0] == 0


def test_context_close_waits_for_servers():
    gb = load_pkg()
    lib = gb.library()
    with gb.Context(0) as c:
        srv = c.server()
        assert lib.gb_close(c._h) == gb.GB_EAGAIN   # refused while the server runs
        assert list(srv.run([98])) == [19]
        srv.close()
        assert srv._h is None
    # a closed context cannot be used for a server
    c2 = gb.Context(0)
    s2 = c2.server()
    c2.close()   # closes its server first
    assert s2._h is None and c2._h is None
