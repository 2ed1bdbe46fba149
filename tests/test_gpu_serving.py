"""Continuous-batching server: each result is the same alone and in the mix (bit-exact),
admission by device memory, queue limits, cancellation and time-outs."""
import sys
import threading

import numpy as np
import pytest

from tpchq01_gpu_amd import q1
from tpchq01_gpu_amd.serving import Server, make_request_stream, run_open_loop

pytestmark = pytest.mark.gpu

SEED = 0x51C00004


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tables(ctx):
    ts = [ctx.generate(q1.COMPRESSED, SEED + k, rows) for k, rows in enumerate([10_000, 1_000_000, 20_000_000])]
    yield ts
    for t in ts:
        t.free()


def test_results_same_alone_and_in_the_mix(ctx, tables):
    """Arrivals faster than one slot serves them (600 requests in 30 ms): the queue fills
    while the slot is busy, so later requests run in batches with others."""
    stream = make_request_stream(tables, rate_per_s=20000, duration_s=0.03, seed=1)
    got = {}

    def collect(i, st, groups):
        got[i] = (st, groups)

    with Server(ctx, max_batch=64, max_inflight=1, batch_window_us=300) as sv:
        done, wall = run_open_loop(sv, stream, collect)
    note(f"served {len(stream)} requests in {wall:.3f}s, batch sizes {sorted({d[2] for d in done})[:10]}")
    assert max(d[2] for d in done) > 1, "requests were never batched together"
    for i, (at, t, b, n, c) in enumerate(stream):
        st, groups = got[i]
        assert st == q1.OK
        assert groups == ctx.run(t, b, n, c), i


def test_concurrent_submitters(ctx, tables):
    rng = np.random.default_rng(2)
    reqs = []
    for _ in range(300):
        t = tables[int(rng.integers(0, 3))]
        n = int(rng.integers(0, t.rows + 1))
        b = int(rng.integers(0, t.rows - n + 1))
        reqs.append((t, b, n, int(rng.integers(729906, 730089))))
    results = [None] * len(reqs)
    with Server(ctx, max_batch=32, max_inflight=3) as sv:
        def worker(lo, hi):
            for i in range(lo, hi):
                t, b, n, c = reqs[i]
                st, tk = sv.submit(t, b, n, c)
                assert st == q1.OK
                results[i] = sv.wait(tk)
        ths = [threading.Thread(target=worker, args=(k * 75, (k + 1) * 75)) for k in range(4)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
    for i, (t, b, n, c) in enumerate(reqs):
        st, groups, _ = results[i]
        assert st == q1.OK and groups == ctx.run(t, b, n, c)


def table_bytes(t):
    return sum(q1.column_bytes(t.layout, t.rows))


def burst(sv, reqs):
    """Submits every request at once; returns (status, groups, stats, t_before, t_after) each,
    the submit call bracketed by host clock readings (seconds)."""
    import time
    subs = []
    for t, b, n, c in reqs:
        t0 = time.perf_counter()
        st, tk = sv.submit(t, b, n, c)
        subs.append((st, tk, t0, time.perf_counter()))
    out = []
    for st, tk, t0, t1 in subs:
        assert st == q1.OK
        st, groups, s = sv.wait(tk)
        out.append((st, groups, s, t0, t1))
    return out


def test_large_requests_run_alone_one_at_a_time(ctx, tables):
    """exclusive_bytes = 1 GiB on a 200M-row table (1.875 GB of columns): its full scans are
    large, its 100M-row halves are not.  Each large request runs in a batch of its own and
    starts only after the large one before it has finished; the small requests that queue
    meanwhile still batch together on the other slot."""
    big = ctx.generate(q1.COMPRESSED, SEED + 7, 200_000_000)
    small = tables[1]
    try:
        reqs = []
        for i in range(6):
            reqs.append((big, 0, big.rows, 729999 - i))
            reqs.append((big, 50_000_000 * (i % 3), 100_000_000, 729980 + i))
            reqs += [(small, 1000 * k, 500_000, 729990 + k) for k in range(12)]
        note(f"burst of {len(reqs)} requests, exclusive_bytes 1 GiB")
        with Server(ctx, max_batch=64, max_inflight=2, exclusive_bytes=1 << 30) as sv:
            res = burst(sv, reqs)
        larges, small_sizes = [], []
        for (t, b, n, c), (st, groups, s, t0, t1) in zip(reqs, res):
            assert st == q1.OK and groups == ctx.run(t, b, n, c)
            if n == big.rows:
                assert s.batch_size == 1
                # host-clock bounds of the request's start and end (the submit call's own
                # duration is the uncertainty)
                larges.append((t0 + s.submit_to_start_us * 1e-6, t1 + s.submit_to_start_us * 1e-6,
                               t0 + s.latency_us * 1e-6, t1 + s.latency_us * 1e-6, s.batch_id))
            elif t is small:
                small_sizes.append(s.batch_size)
        larges.sort(key=lambda x: x[4])
        for x, y in zip(larges, larges[1:]):
            assert y[1] >= x[2], "a large request started before the previous one finished"
        note(f"small-request batch sizes {sorted(set(small_sizes))}")
        assert max(small_sizes) > 1
        # exclusive_bytes = 2**64 - 1: no request is large, full scans batch together too
        reqs = [(big, 0, big.rows, 729999 - i) for i in range(8)]
        with Server(ctx, max_batch=64, max_inflight=1, exclusive_bytes=2**64 - 1) as sv:
            res = burst(sv, reqs)
        assert max(r[2].batch_size for r in res) > 1
        for (t, b, n, c), (st, groups, *_) in zip(reqs, res):
            assert st == q1.OK and groups == ctx.run(t, b, n, c)
    finally:
        big.free()


def test_admission_by_memory_budget(ctx, tables):
    """In-flight work pins its tables' columns (once per table) and 480 B of descriptor and
    result per request: a budget of one table plus four requests' state keeps at most four
    requests on the device at once."""
    per_request = 96 + 384   # descriptor + result table bytes on the device
    with Server(ctx, max_batch=256, max_inflight=2, memory_budget=table_bytes(tables[1]) + 4 * per_request) as sv:
        tickets = []
        for i in range(60):
            st, tk = sv.submit(tables[1], i * 1000, 500_000, 729999)
            assert st == q1.OK
            tickets.append(tk)
        sizes = []
        for tk in tickets:
            st, groups, s = sv.wait(tk)
            assert st == q1.OK
            sizes.append(s.batch_size)
    assert max(sizes) <= 4
    # a budget below one request's state can never admit anything
    with Server(ctx, memory_budget=100) as sv:
        st, _ = sv.submit(tables[0], 0, 10, 729999)
        assert st == q1.ENOMEM


def test_queue_full_cancel_and_timeout(ctx, tables):
    with Server(ctx, max_batch=4, max_inflight=1, queue_capacity=8, batch_window_us=0) as sv:
        statuses, tickets = [], []
        for i in range(200):
            st, tk = sv.submit(tables[2], 0, tables[2].rows, 729999)
            statuses.append(st)
            if st == q1.OK:
                tickets.append(tk)
        assert q1.EBUSY in statuses
        canceled = 0
        for tk in tickets[::-1]:
            c = sv.cancel(tk)
            assert c in (q1.OK, q1.EBUSY)
            canceled += c == q1.OK
        for tk in tickets:
            st, groups, _ = sv.wait(tk, timeout_us=10_000_000)
            assert st in (q1.OK, q1.ECANCELED)
        assert canceled >= 1
        st, tk = sv.submit(tables[2], 0, tables[2].rows, 729999)
        st0, _, _ = sv.wait(tk, timeout_us=0)
        assert st0 in (q1.ETIMEOUT, q1.OK)
        if st0 == q1.ETIMEOUT:
            assert sv.wait(tk)[0] == q1.OK
        assert sv.wait(123456789)[0] == q1.EINVAL


def test_out_of_domain_table_is_served(ctx, oracle):
    import pinning
    rows = pinning.random_rows(np.random.default_rng(3), 5000, tpch=False)
    cols = pinning.ref_columns(rows)
    t = ctx.upload(q1.REFERENCE, cols, len(rows))
    with Server(ctx) as sv:
        st, tk = sv.submit(t, 0, len(rows), 729999)
        st, groups, _ = sv.wait(tk)
    want, _ = oracle.naive(cols, 0, len(rows), 729999)
    assert st == q1.OK
    assert [dict(sum_qty=g.sum_qty, sum_base_price=g.sum_base_price, sum_disc_price=g.sum_disc_price,
                 sum_charge=g.sum_charge, sum_disc=g.sum_disc, count=g.count) for g in groups] == want
