"""The HIP path against the oracle at the sizes of every measured workload, bit-exact,
through the C ABI:

  * 6,000,000 rows: every layout x placement x cutoff vs KernelNaive (and KernelHyPer
    raising no overflow);
  * 60,000,000 uncompressed rows: every placement vs KernelNaive;
  * 600,000,000 device-generated compressed rows: every placement, the pushdown table,
    q1_enqueue with and without the domain flag, and the CPU/GPU split at the reference's
    fraction 0.5, vs KernelNaive over the same seeded rows regenerated on the host in
    chunks;
  * 4,500,000,017 device-generated rows in the compressed, uncompressed and reference
    layouts (42, 90 and 171 GB): counts, offsets and batch requests past 2^32, against
    chunked runs and against KernelNaive on the ranges around row 2^32 and at the end;
  * the 256-request batch (uneven log-uniform sizes, own cutoffs): every request vs
    KernelNaive on its own rows;
  * the serving mix (8 resident tables of 1M..600M rows, Poisson arrivals at 2000 req/s):
    every served result vs the same request run alone, and 200 sampled results vs the
    oracle's compressed kernel.

The oracle is the KernelNaive restatement (cpu/kernel_naive.hpp:12-28) in oracle/; large
row counts are checked chunk by chunk (morsel-parallel restatements, exact sums of the
chunks' results in Python integers).
"""
import os
import sys

import numpy as np
import pytest

from tpchq01_gpu_amd import q1, workloads

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

THREADS = max(1, min(16, os.cpu_count() or 4))
CUTOFFS = [729999, 729906, 730088]   # 1998-09-02, 1998-06-01, 1998-11-30
PLACEMENTS = [q1.IN_REGISTERS, q1.IN_REGISTERS_PER_THREAD, q1.LOCAL_MEM, q1.PER_THREAD_SHARED_MEM, q1.GLOBAL]
FIELDS = ("sum_qty", "sum_base_price", "sum_disc_price", "sum_charge", "sum_disc", "count")


def as_dicts(groups):
    return [dict(sum_qty=g.sum_qty, sum_base_price=g.sum_base_price, sum_disc_price=g.sum_disc_price,
                 sum_charge=g.sum_charge, sum_disc=g.sum_disc, count=g.count) for g in groups]


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def add_results(a, b):
    return [{f: x[f] + y[f] for f in FIELDS} for x, y in zip(a, b)]


def oracle_seeded(oracle, seed, begin, count, cutoffs, chunk=50_000_000):
    """KernelNaive (morsel-parallel restatement) over rows [begin, begin + count) of the
    seeded generator, regenerated on the host in reference-layout chunks; one result per
    cutoff, the chunks' results summed exactly."""
    assert begin % 32 == 0 and chunk % 32 == 0
    total = {c: [dict.fromkeys(FIELDS, 0) for _ in range(6)] for c in cutoffs}
    done = 0
    while done < count:
        n = min(chunk, count - done)
        ref = q1.generate_host(q1.REFERENCE, seed, n, begin + done, THREADS)
        for c in cutoffs:
            got, unexpected = oracle.naive_parallel(ref, 0, n, c, THREADS)
            assert unexpected == 0
            total[c] = add_results(total[c], got)
        done += n
        del ref
    return total


def test_cpu_reference_case_6M_rows(ctx, oracle):
    """6,000,000 seeded rows: all four layouts x five placements x three cutoffs."""
    rows, seed = 6_000_000, workloads.SAMPLE_SEED
    h = {lay: q1.generate_host(lay, seed, rows, 0, THREADS) for lay in (q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED)}
    tables = {lay: ctx.upload(lay, h[lay], rows) for lay in h}
    try:
        for cutoff in CUTOFFS:
            want, unexpected = oracle.naive(h[q1.REFERENCE], 0, rows, cutoff)
            assert unexpected == 0
            hyper, overflow = oracle.hyper(h[q1.REFERENCE], 0, rows, cutoff)
            assert overflow == 0 and hyper == want
            pd = ctx.build_filter(tables[q1.COMPRESSED], cutoff)
            for layout, t in list(tables.items()) + [(q1.COMPRESSED_PUSHDOWN, pd)]:
                for p in PLACEMENTS + [q1.AUTO]:
                    note(f"case 6M rows cutoff={cutoff} layout={q1.LAYOUT_NAMES[layout]} placement={p}")
                    assert as_dicts(ctx.run(t, 0, rows, cutoff, p)) == want, (cutoff, layout, p)
            pd.free()
    finally:
        for t in tables.values():
            t.free()


def test_uncompressed_60M_rows_every_placement(ctx, oracle):
    """60,000,000 device-resident uncompressed rows (20 B/row) through every placement."""
    rows, seed = 60_000_000, workloads.SAMPLE_SEED
    t = ctx.generate(q1.UNCOMPRESSED, seed, rows)
    try:
        want = oracle_seeded(oracle, seed, 0, rows, [729999, 729906])
        for cutoff, w in want.items():
            for p in PLACEMENTS:
                note(f"case 60M uncompressed cutoff={cutoff} placement={q1.PLACEMENT_NAMES[p]}")
                assert as_dicts(ctx.run(t, 0, rows, cutoff, p)) == w, (cutoff, p)
    finally:
        t.free()


def test_scan_600M_compressed_rows(ctx, oracle):
    """600,000,000 device-generated compressed rows: every placement at its tuned launch,
    the device-built pushdown table, the drop-in q1_enqueue with and without the domain
    flag, and the CPU/GPU split at fraction 0.5 with the host's copy of the rows."""
    import ctypes as C

    import torch
    rows, seed = workloads.SCAN_ROWS, workloads.SCAN_SEED
    note("case 600M: oracle over the host-regenerated rows")
    want = oracle_seeded(oracle, seed, 0, rows, [729999, 730088])
    t = ctx.generate(q1.COMPRESSED, seed, rows)
    try:
        for cutoff, w in want.items():
            for p in PLACEMENTS + [q1.AUTO]:
                note(f"case 600M compressed cutoff={cutoff} placement={p}")
                assert as_dicts(ctx.run(t, 0, rows, cutoff, p)) == w, (cutoff, p)
        w = want[729999]
        pd = ctx.build_filter(t, 729999)
        for p in (q1.AUTO, q1.IN_REGISTERS):
            note(f"case 600M pushdown placement={p}")
            assert as_dicts(ctx.run(pd, 0, rows, 0, p)) == w
        pd.free()
        cols = t.device_columns()
        agg = torch.zeros(48, dtype=torch.int64, device="cuda")
        for flags in (q1.LAUNCH_ASSUME_TPCH_DOMAIN, 0):
            note(f"case 600M q1_enqueue flags={flags}")
            agg.zero_()
            torch.cuda.synchronize()
            launch = q1.make_launch(flags=flags)
            ctx.check(q1.lib.q1_enqueue(ctx.handle, q1.PER_THREAD_SHARED_MEM, q1.COMPRESSED, C.byref(launch),
                                        C.c_void_p(agg.data_ptr()), C.byref(cols), 0, rows, 729999, None),
                      "q1_enqueue")
            ctx.synchronize()
            host = q1.DeviceAggregates()
            arr = agg.cpu().numpy()
            C.memmove(C.byref(host), arr.ctypes.data, C.sizeof(host))
            r = q1.Result()
            q1.lib.q1_aggregates_to_result(C.byref(host), C.byref(r))
            assert as_dicts(q1.result_to_list(r)) == w, flags
        note("case 600M split 0.5")
        hcols = q1.generate_host(q1.COMPRESSED, seed, rows, 0, THREADS)
        got, cpu_ms, gpu_ms = ctx.run_split(t, hcols, rows, 0.5, 729999, q1.AUTO, None, 1 << 21, THREADS)
        assert as_dicts(got) == w
        note(f"case 600M split: gpu {gpu_ms:.3f} ms, cpu {cpu_ms:.3f} ms")
    finally:
        t.free()


@pytest.mark.parametrize("layout", [q1.COMPRESSED, q1.UNCOMPRESSED, q1.REFERENCE],
                         ids=["compressed_42GB", "uncompressed_90GB", "reference_171GB"])
def test_scan_beyond_2_32_rows(ctx, oracle, layout):
    """4,500,000,017 device-generated rows in each layout (42, 90 and 171 GB of columns): row
    counts and offsets past 2^32 (the reference counts rows in 32 bits, this ABI in 64), and
    byte offsets of the 8-byte columns past 2^35.
      * at the generator's last ship date (1998-12-01) every row passes: in every
        placement the groups' counts add up to the row count;
      * at 1998-09-02 every placement, the pushdown table (compressed) and a batch request
        over the whole range equal the exact sum of runs over 2^30-row chunks;
      * 2^25 rows straddling row 2^32, the 12 rows around it and the ragged last ~1M rows
        equal KernelNaive over the same rows regenerated on the host, in every placement
        and as a batch request."""
    rows, seed, cutoff, every_row = 4_500_000_017, workloads.SCAN_SEED, 729999, 730089
    name = q1.LAYOUT_NAMES[layout]
    t = ctx.generate(layout, seed, rows)
    try:
        for p in PLACEMENTS + [q1.AUTO]:
            note(f"case 4.5G {name} rows all pass placement={p}")
            got = as_dicts(ctx.run(t, 0, rows, every_row, p))
            assert sum(g["count"] for g in got) == rows, p
        note(f"case 4.5G {name} rows in 2^30-row chunks")
        chunked = [dict.fromkeys(FIELDS, 0) for _ in range(6)]
        for b in range(0, rows, 1 << 30):
            chunked = add_results(chunked, as_dicts(ctx.run(t, b, min(1 << 30, rows - b), cutoff, q1.AUTO)))
        for p in PLACEMENTS + [q1.AUTO]:
            note(f"case 4.5G {name} rows whole range placement={p}")
            assert as_dicts(ctx.run(t, 0, rows, cutoff, p)) == chunked, p
        if layout == q1.COMPRESSED:
            note(f"case 4.5G {name} rows pushdown")
            pd = ctx.build_filter(t, cutoff)
            try:
                assert as_dicts(ctx.run(pd, 0, rows, 0, q1.AUTO)) == chunked
            finally:
                pd.free()

        straddle = (2**32 - 2**24, 2**25)
        tail_begin = (rows - 1_000_000) // 32 * 32
        ranges = [straddle, (tail_begin, rows - tail_begin)]
        for begin, count in ranges:
            note(f"case 4.5G {name} rows oracle over [{begin}, {begin + count})")
            want = oracle_seeded(oracle, seed, begin, count, [cutoff])[cutoff]
            for p in PLACEMENTS + [q1.AUTO]:
                assert as_dicts(ctx.run(t, begin, count, cutoff, p)) == want, (begin, p)
            if begin == straddle[0]:
                want_straddle = want
        note(f"case 4.5G {name} rows: the 12 rows around row 2^32")
        ref = q1.generate_host(q1.REFERENCE, seed, 64, 2**32 - 32, 1)
        want, unexpected = oracle.naive_parallel(ref, 27, 39, every_row, 1)
        assert unexpected == 0 and sum(g["count"] for g in want) == 12
        for p in PLACEMENTS + [q1.AUTO]:
            assert as_dicts(ctx.run(t, 2**32 - 5, 12, every_row, p)) == want, p

        note(f"case 4.5G {name} rows batch")
        reqs = (q1.Request * 2)(q1.Request(t.handle, 0, rows, cutoff, 0),
                                q1.Request(t.handle, straddle[0], straddle[1], cutoff, 0))
        outs = (q1.Result * 2)()
        ctx.check(q1.lib.q1_run_batch(ctx.handle, reqs, 2, None, outs), "q1_run_batch")
        assert as_dicts(q1.result_to_list(outs[0])) == chunked
        assert as_dicts(q1.result_to_list(outs[1])) == want_straddle
    finally:
        t.free()


def test_batch_256_uneven_requests(ctx, oracle):
    """The 256-request batch (request r: its own table of seed 0x51C00003 ^ r, log-uniform
    1e4..2e7 rows, its own cutoff) in one q1_run_batch: every request vs KernelNaive over
    its rows."""
    spec = workloads.batch_spec(256)
    tables = [ctx.generate(q1.COMPRESSED, s, n) for s, n, _ in spec]
    try:
        n = len(spec)
        reqs = (q1.Request * n)(*[q1.Request(tables[r].handle, 0, spec[r][1], spec[r][2], 0) for r in range(n)])
        outs = (q1.Result * n)()
        note(f"case batch of {n} requests, {sum(x[1] for x in spec)} rows")
        for rep in range(2):   # the first call captures the graph, the second replays it
            ctx.check(q1.lib.q1_run_batch(ctx.handle, reqs, n, None, outs), "q1_run_batch")
            got = [as_dicts(q1.result_to_list(outs[r])) for r in range(n)]
            if rep == 0:
                first = got
            assert got == first
        for r, (s, rows, cutoff) in enumerate(spec):
            ref = q1.generate_host(q1.REFERENCE, s, rows, 0, THREADS)
            want, _ = oracle.naive_parallel(ref, 0, rows, cutoff, THREADS)
            assert got[r] == want, (r, s, rows, cutoff)
    finally:
        for t in tables:
            t.free()


def test_serving_mix_2000_req_s(ctx, oracle):
    """The serving mix: 8 resident compressed tables of 1M..600M rows, Poisson arrivals at
    2000 req/s for one second through the continuous-batching server.  Every served result
    equals the request run alone; 200 sampled results equal the oracle's compressed kernel
    on the host's copy of the table."""
    from tpchq01_gpu_amd.serving import Server, make_request_stream, run_open_loop
    spec = workloads.serving_spec()
    tables = [ctx.generate(q1.COMPRESSED, s, n) for s, n in spec]
    try:
        stream = make_request_stream(tables, rate_per_s=2000, duration_s=1.0, seed=9)
        got = {}

        def collect(i, st, groups):
            got[i] = (st, groups)

        with Server(ctx, max_batch=256, max_inflight=2, batch_window_us=200) as sv:
            done, wall = run_open_loop(sv, stream, collect)
        note(f"case serving: {len(stream)} requests in {wall:.3f} s, mean batch "
             f"{np.mean([d[2] for d in done]):.2f}")
        for i, (at, t, b, n, c) in enumerate(stream):
            st, groups = got[i]
            assert st == q1.OK, i
            assert groups == ctx.run(t, b, n, c), i
        # admission by device memory at this scale: a budget that holds every table but the
        # 600M-row one refuses requests on that table at submit and serves the others
        sizes = [sum(q1.column_bytes(t.layout, t.rows)) for t in tables]
        with Server(ctx, max_batch=256, max_inflight=2, memory_budget=sum(sizes[:-1]) + (1 << 20)) as sv:
            for k, t in enumerate(tables):
                st, tk = sv.submit(t, 0, t.rows, 729999)
                if k == len(tables) - 1:
                    assert st == q1.ENOMEM and sizes[k] > sum(sizes[:-1])
                else:
                    assert st == q1.OK
                    st, groups, _ = sv.wait(tk)
                    assert st == q1.OK and groups == ctx.run(t, 0, t.rows, 729999), k
        rng = np.random.default_rng(0x51C0)
        sample = sorted(rng.choice(len(stream), 200, replace=False))
        index = {t.handle: k for k, t in enumerate(tables)}
        by_table = {}
        for i in sample:
            by_table.setdefault(index[stream[i][1].handle], []).append(i)
        for k, idx in sorted(by_table.items()):
            seed, rows = spec[k]
            note(f"case serving oracle: table {k} ({rows} rows), {len(idx)} requests")
            host = q1.generate_host(q1.COMPRESSED, seed, rows, 0, THREADS)
            for i in idx:
                _, _, b, n, c = stream[i]
                want = oracle.compressed_parallel(host, b, b + n, c, THREADS)
                for g in want:
                    g["sum_qty"] *= 100
                assert as_dicts(got[i][1]) == want, (i, k, b, n, c)
            del host
    finally:
        for t in tables:
            t.free()
