"""The drop-in kernel launch (q1_enqueue), wrapped device columns, batching and
errors, through the C ABI on the GPU."""
import ctypes as C
import sys

import numpy as np
import pytest

from tpchq01_gpu_amd import q1

pytestmark = pytest.mark.gpu

SEED = 0x51C00003


def as_dicts(groups):
    return [dict(sum_qty=g.sum_qty, sum_base_price=g.sum_base_price, sum_disc_price=g.sum_disc_price,
                 sum_charge=g.sum_charge, sum_disc=g.sum_disc, count=g.count) for g in groups]


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_aggregates_to_result(t):
    host = q1.DeviceAggregates()
    arr = t.cpu().numpy()   # keep the host copy alive while it is read
    C.memmove(C.byref(host), arr.ctypes.data, C.sizeof(host))
    r = q1.Result()
    q1.lib.q1_aggregates_to_result(C.byref(host), C.byref(r))
    return q1.result_to_list(r)


def test_enqueue_composes_by_addition(ctx, oracle):
    """Launches over disjoint row ranges add into one caller-zeroed device table, like the
    reference's per-chunk launches (src/main.cu:553-558, 657-718)."""
    import torch
    rows = 2_000_011
    ref = q1.generate_host(q1.REFERENCE, SEED, rows)
    comp = q1.generate_host(q1.COMPRESSED, SEED, rows)
    want, _ = oracle.naive(ref, 0, rows, 729999)
    t = ctx.upload(q1.COMPRESSED, comp, rows)
    cols = t.device_columns()
    for p in range(5):
        agg = torch.zeros(48, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        cuts = [0, 17, 1 << 20, (1 << 20) + 33, rows]
        for b, e in zip(cuts, cuts[1:]):
            l = q1.make_launch(flags=q1.LAUNCH_ASSUME_TPCH_DOMAIN)
            ctx.check(q1.lib.q1_enqueue(ctx.handle, p, q1.COMPRESSED, C.byref(l), C.c_void_p(agg.data_ptr()),
                                        C.byref(cols), b, e - b, 729999, None), "q1_enqueue")
        ctx.synchronize()
        assert as_dicts(device_aggregates_to_result(agg)) == want, p
        # same without the domain assertion (64-bit kernels)
        agg.zero_()
        torch.cuda.synchronize()
        ctx.check(q1.lib.q1_enqueue(ctx.handle, p, q1.COMPRESSED, None, C.c_void_p(agg.data_ptr()),
                                    C.byref(cols), 0, rows, 729999, None), "q1_enqueue")
        ctx.synchronize()
        assert as_dicts(device_aggregates_to_result(agg)) == want, p


def test_enqueue_other_stream_and_context_stream_agree(ctx, oracle):
    """Launches on the context stream merge large grids through the context's spread
    tables, launches on another stream through that stream's own.  Both add into the
    caller's table, and back-to-back launches on each keep adding exactly."""
    import torch
    rows = 3_000_017
    ref = q1.generate_host(q1.REFERENCE, SEED + 2, rows)
    comp = q1.generate_host(q1.COMPRESSED, SEED + 2, rows)
    want, _ = oracle.naive(ref, 0, rows, 729999)
    t = ctx.upload(q1.COMPRESSED, comp, rows)
    cols = t.device_columns()
    other = torch.cuda.Stream()
    for p in (q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM):
        for stream in (None, other.cuda_stream):
            agg = torch.zeros(48, dtype=torch.int64, device="cuda")
            torch.cuda.synchronize()
            l = q1.make_launch(64, 16, 0, q1.LAUNCH_ASSUME_TPCH_DOMAIN)   # 2930 blocks
            for _ in range(3):
                ctx.check(q1.lib.q1_enqueue(ctx.handle, p, q1.COMPRESSED, C.byref(l), C.c_void_p(agg.data_ptr()),
                                            C.byref(cols), 0, rows, 729999, C.c_void_p(stream)), "q1_enqueue")
            ctx.synchronize()
            torch.cuda.synchronize()
            got = device_aggregates_to_result(agg)
            for g, w in zip(got, want):
                assert (g.count, g.sum_qty, g.sum_charge, g.sum_disc_price, g.sum_base_price, g.sum_disc) == (
                    3 * w["count"], 3 * w["sum_qty"], 3 * w["sum_charge"], 3 * w["sum_disc_price"],
                    3 * w["sum_base_price"], 3 * w["sum_disc"]), (p, stream)
    t.free()


def _enqueue(ctx, p, cols, agg, rows, stream, launch):
    ctx.check(q1.lib.q1_enqueue(ctx.handle, p, q1.COMPRESSED, C.byref(launch), C.c_void_p(agg.data_ptr()),
                                C.byref(cols), 0, rows, 729999, C.c_void_p(stream)), "q1_enqueue")


def _assert_times(agg, want, k, what):
    got = device_aggregates_to_result(agg)
    for g, w in zip(got, want):
        assert (g.count, g.sum_qty, g.sum_charge, g.sum_disc_price, g.sum_base_price, g.sum_disc) == (
            k * w["count"], k * w["sum_qty"], k * w["sum_charge"], k * w["sum_disc_price"],
            k * w["sum_base_price"], k * w["sum_disc"]), what


def test_enqueue_on_many_streams_and_a_recreated_stream(ctx, oracle):
    """Large launches on 20 caller streams (the first 16 get spread tables of their own,
    the rest merge directly), two streams adding into one table at once, and a stream
    destroyed with its launch still queued and created again (possibly at the same
    address, so its tables are reused after the event of their last use): every table
    holds exactly the launches added into it."""
    import torch
    rows = 1_500_007
    ref = q1.generate_host(q1.REFERENCE, SEED + 5, rows)
    comp = q1.generate_host(q1.COMPRESSED, SEED + 5, rows)
    want, _ = oracle.naive(ref, 0, rows, 729999)
    t = ctx.upload(q1.COMPRESSED, comp, rows)
    cols = t.device_columns()
    l = q1.make_launch(64, 16, 0, q1.LAUNCH_ASSUME_TPCH_DOMAIN)   # 1465 blocks
    streams = [torch.cuda.Stream() for _ in range(20)]
    aggs = [torch.zeros(48, dtype=torch.int64, device="cuda") for _ in streams]
    shared = torch.zeros(48, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    note("case 20 streams")
    for rep in range(2):
        for s, a in zip(streams, aggs):
            _enqueue(ctx, q1.PER_THREAD_SHARED_MEM, cols, a, rows, s.cuda_stream, l)
        _enqueue(ctx, q1.IN_REGISTERS, cols, shared, rows, streams[0].cuda_stream, l)
        _enqueue(ctx, q1.IN_REGISTERS, cols, shared, rows, streams[19].cuda_stream, l)
    torch.cuda.synchronize()
    for i, a in enumerate(aggs):
        _assert_times(a, want, 2, i)
    _assert_times(shared, want, 4, "shared")
    t.free()
    note("case recreated stream")
    # the HIP runtime this process already uses (the library's and torch's: one copy)
    paths = sorted({line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line})
    assert len(paths) == 1, paths
    hip = C.CDLL(paths[0])
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    with q1.Context(0) as c2:      # a context whose stream tables are not used up yet
        t2 = c2.upload(q1.COMPRESSED, comp, rows)
        cols2 = t2.device_columns()
        agg = torch.zeros(48, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        for rep in range(3):
            h = C.c_void_p()
            assert hip.hipStreamCreate(C.byref(h)) == 0
            _enqueue(c2, q1.PER_THREAD_SHARED_MEM, cols2, agg, rows, h.value, l)
            assert hip.hipStreamDestroy(h) == 0
        c2.synchronize()
        torch.cuda.synchronize()
        _assert_times(agg, want, 3, "recreated")
        t2.free()


def test_enqueue_under_graph_capture(ctx, oracle):
    """q1_enqueue captured into a graph on a caller's stream (it merges directly there and
    allocates nothing): every replay adds exactly one result, with eager launches between
    the replays."""
    import torch
    rows = 1_500_007
    ref = q1.generate_host(q1.REFERENCE, SEED + 6, rows)
    comp = q1.generate_host(q1.COMPRESSED, SEED + 6, rows)
    want, _ = oracle.naive(ref, 0, rows, 729999)
    t = ctx.upload(q1.COMPRESSED, comp, rows)
    cols = t.device_columns()
    l = q1.make_launch(64, 16, 0, q1.LAUNCH_ASSUME_TPCH_DOMAIN)
    agg = torch.zeros(48, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    note("case capture")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        _enqueue(ctx, q1.PER_THREAD_SHARED_MEM, cols, agg, rows, s.cuda_stream, l)
    torch.cuda.synchronize()
    _assert_times(agg, want, 0, "captured, not run")
    k = 0
    for rep in range(3):
        g.replay()
        torch.cuda.synchronize()
        k += 1
        _assert_times(agg, want, k, ("replay", rep))
        _enqueue(ctx, q1.IN_REGISTERS, cols, agg, rows, s.cuda_stream, l)   # eager, between replays
        torch.cuda.synchronize()
        k += 1
        _assert_times(agg, want, k, ("eager", rep))
    del g
    t.free()


def test_wrap_torch_columns(ctx, oracle):
    import torch
    rows = 123_457
    ref = q1.generate_host(q1.REFERENCE, SEED + 1, rows)
    unc = q1.generate_host(q1.UNCOMPRESSED, SEED + 1, rows)
    want, _ = oracle.naive(ref, 0, rows, 729906)
    dev = {k: torch.from_numpy(v.copy()).to("cuda") for k, v in unc.items()}
    torch.cuda.synchronize()
    c = q1.Columns()
    for k, v in dev.items():
        setattr(c, k, v.data_ptr())
    t = ctx.wrap(q1.UNCOMPRESSED, c, rows)
    for p in range(5):
        assert as_dicts(ctx.run(t, 0, rows, 729906, p)) == want
    t.free()
    # misaligned column pointer is refused
    c.quantity = dev["quantity"].data_ptr() + 4
    with pytest.raises(q1.Q1Error):
        ctx.wrap(q1.UNCOMPRESSED, c, rows - 1)


def test_table_outliving_its_context():
    """Tables still alive when their context shuts down lose their device memory; their
    handles stay valid for q1_table_free (what Python's garbage collector may do last)."""
    c = q1.Context(0)
    t = c.generate(q1.COMPRESSED, SEED, 10_000)
    pd = c.build_filter(t, 729999)
    u = c.generate(q1.REFERENCE, SEED, 1000)
    u.free()
    c.close()
    pd.free()
    t.free()


def test_errors(ctx):
    t = ctx.generate(q1.COMPRESSED, SEED, 1000)
    with pytest.raises(q1.Q1Error):
        ctx.run(t, 900, 200)                      # range past the end
    with pytest.raises(q1.Q1Error):
        ctx.run(t, 0, 1000, launch=q1.make_launch(100, 1024))   # not 64, 128 or 256
    with pytest.raises(q1.Q1Error):
        ctx.run(t, 0, 1000, launch=q1.make_launch(192, 1024))
    with pytest.raises(q1.Q1Error):
        ctx.run(t, 0, 1000, launch=q1.make_launch(256, 1000))   # not a multiple of 16
    with pytest.raises(q1.Q1Error):
        ctx.run(t, 0, 1000, launch=q1.make_launch(256, 1 << 15))  # above 16384
    with pytest.raises(q1.Q1Error):
        ctx.build_filter(ctx.generate(q1.REFERENCE, SEED, 10), 729999)
    assert ctx.run(t, 1000, 0)[0].count == 0     # empty range


def test_device_info(ctx):
    d = ctx.device_info()
    note("device", d.name, d.arch, d.compute_units, d.total_memory)
    assert d.compute_units > 0 and d.wavefront_size == 64
    assert b"gfx950" in d.arch


def make_requests(ctx, tables, rng, n):
    reqs, meta = [], []
    for i in range(n):
        t = tables[int(rng.integers(0, len(tables)))]
        size = int(np.exp(rng.uniform(np.log(1), np.log(t.rows))))
        begin = int(rng.integers(0, t.rows - size + 1))
        cutoff = int(rng.integers(729906, 730089))
        reqs.append(q1.Request(t.handle, begin, size, cutoff, 0))
        meta.append((t, begin, size, cutoff))
    return reqs, meta


def test_batch_matches_single_runs_across_graph_replays(ctx):
    rng = np.random.default_rng(1)
    tables = [ctx.generate(q1.COMPRESSED, SEED + k, rows) for k, rows in enumerate([1000, 65_536, 1_000_003, 3_000_000])]
    for rep, n in enumerate([1, 7, 16, 16, 100, 256, 3, 256]):
        reqs, meta = make_requests(ctx, tables, rng, n)
        arr = (q1.Request * n)(*reqs)
        outs = (q1.Result * n)()
        note(f"case batch replay {rep} n={n}")
        ctx.check(q1.lib.q1_run_batch(ctx.handle, arr, n, None, outs), "q1_run_batch")
        # eager work between replays, then every request alone
        for i, (t, b, s, c) in enumerate(meta):
            assert q1.result_to_list(outs[i]) == ctx.run(t, b, s, c, q1.IN_REGISTERS), (rep, i)


def test_batch_pushdown_and_mixed_layouts(ctx, oracle):
    rows = 200_000
    comp = ctx.generate(q1.COMPRESSED, SEED + 5, rows)
    pd = ctx.build_filter(comp, 729950)
    ref = ctx.generate(q1.REFERENCE, SEED + 5, rows)
    reqs = [q1.Request(pd.handle, 5, rows - 10, 0, 0), q1.Request(pd.handle, 0, 77, 0, 0)]
    outs = (q1.Result * 2)()
    ctx.check(q1.lib.q1_run_batch(ctx.handle, (q1.Request * 2)(*reqs), 2, None, outs), "q1_run_batch")
    assert q1.result_to_list(outs[0]) == ctx.run(comp, 5, rows - 10, 729950)
    assert q1.result_to_list(outs[1]) == ctx.run(comp, 0, 77, 729950)
    # a mix of layouts runs request by request, same answers
    reqs = [q1.Request(comp.handle, 0, rows, 729999, 0), q1.Request(ref.handle, 3, 1000, 729999, 0)]
    outs = (q1.Result * 2)()
    ctx.check(q1.lib.q1_run_batch(ctx.handle, (q1.Request * 2)(*reqs), 2, None, outs), "q1_run_batch")
    assert q1.result_to_list(outs[0]) == ctx.run(comp, 0, rows, 729999)
    assert q1.result_to_list(outs[1]) == ctx.run(ref, 3, 1000, 729999)


@pytest.mark.parametrize("layout", [q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED])
def test_batch_every_layout_and_batch_placement(layout):
    """The batched kernel for each layout with both of its placements (the register table,
    and the LDS table that reads the columns in small streaming tiles), on ragged ranges,
    at the tuned rows per lane and at 16 (many blocks per request): every result equals
    the request run alone."""
    rng = np.random.default_rng(7 + layout)
    with q1.Context(0) as c:
        tables = [c.generate(layout, SEED + 11 + k, rows) for k, rows in enumerate([1001, 70_003, 1_000_001])]
        for p in (q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM):
            c.set_best_placement(layout, p)
            for n, tpt in ((1, 0), (33, 0), (200, 0), (60, 16)):
                reqs, meta = make_requests(c, tables, rng, n)
                outs = (q1.Result * n)()
                note(f"case batch layout={q1.LAYOUT_NAMES[layout]} placement={p} n={n} tpt={tpt}")
                c.check(q1.lib.q1_run_batch(c.handle, (q1.Request * n)(*reqs), n,
                                            C.byref(q1.make_launch(0, tpt, 0)), outs), "q1_run_batch")
                assert c.last_timing().placement == p
                for i, (t, b, s, cut) in enumerate(meta):
                    assert q1.result_to_list(outs[i]) == c.run(t, b, s, cut, q1.IN_REGISTERS), (p, n, tpt, i)
        for t in tables:
            t.free()


def test_pinned_columns_stream_and_split(ctx):
    """Pinned host columns filled in place by the host generator: streamed to the device and
    split between the device table and host threads, both equal the device-resident run."""
    rows = 1_000_003
    t = ctx.generate(q1.COMPRESSED, SEED + 21, rows)
    want = ctx.run(t, 0, rows, 729999)
    cols = ctx.pinned_columns(q1.COMPRESSED, rows)
    try:
        assert set(cols) == {"ship_date", "quantity", "extended_price", "discount", "tax", "return_flag",
                             "line_status"}
        q1.fill_host(q1.COMPRESSED, SEED + 21, cols, rows, 0, 4)
        note("case pinned streaming")
        got = ctx.run_streaming(q1.COMPRESSED, cols, rows, 729999, q1.AUTO,
                                q1.make_launch(flags=q1.LAUNCH_ASSUME_TPCH_DOMAIN), 1 << 18, 3)
        assert got == want
        for frac in (0.0, 0.003, 0.5, 1.0):
            note(f"case pinned split {frac}")
            got, cpu_ms, gpu_ms = ctx.run_split(t, cols, rows, frac, 729999, q1.AUTO, None, 1 << 16, 4)
            assert got == want, frac
    finally:
        ctx.free_pinned_columns(cols)
        t.free()
    assert cols == {}


def test_batch_small_grids_span_many_requests(ctx):
    """Grids of 1, 2 and 3 blocks make every block work through many requests' blocks,
    one after another (256 requests of about 65 tiles, aligned and unaligned starts):
    every request must still be counted whole, as with the default grid."""
    t = ctx.generate(q1.COMPRESSED, SEED + 31, 70_000)
    for rows, begins in ((1040, [0] * 256), (1040, [16 * (i % 7) for i in range(256)]),
                         (1037, [3 + (i % 5) for i in range(256)])):
        reqs = [q1.Request(t.handle, b, rows, 729999 - (i % 3), 0) for i, b in enumerate(begins)]
        n = len(reqs)
        want = [ctx.run(t, b, rows, 729999 - (i % 3), q1.IN_REGISTERS) for i, b in enumerate(begins)]
        for blocks in (1, 2, 3, 0):
            outs = (q1.Result * n)()
            note(f"case batch search rows={rows} blocks={blocks}")
            ctx.check(q1.lib.q1_run_batch(ctx.handle, (q1.Request * n)(*reqs), n,
                                          C.byref(q1.make_launch(0, 0, blocks)), outs), "q1_run_batch")
            for i in range(n):
                assert q1.result_to_list(outs[i]) == want[i], (rows, blocks, i)
    t.free()


def test_batch_beyond_the_spread_requests(ctx):
    """A batch of 1500 requests: the first 1024 merge through their spread tables, the rest
    straight into their results (the graph then zero-fills and copies the device results).
    Requests of both kinds share entries' neighbourhoods, some are empty, one is large;
    every result equals the request run alone, on two calls in a row (the spread tables
    must come back clean from the first)."""
    rng = np.random.default_rng(SEED + 40)
    t = ctx.generate(q1.COMPRESSED, SEED + 41, 3_000_017)
    n = 1500
    sizes = rng.integers(0, 40_000, n)
    sizes[[0, 700, 1023, 1024, 1499]] = [0, 0, 2_500_000, 0, 1]
    begins = [int(rng.integers(0, t.rows - s + 1)) for s in sizes]
    cutoffs = [729999 - int(rng.integers(0, 120)) for _ in range(n)]
    reqs = (q1.Request * n)(*[q1.Request(t.handle, b, int(s), c, 0) for b, s, c in zip(begins, sizes, cutoffs)])
    want = [ctx.run(t, b, int(s), c, q1.IN_REGISTERS) for b, s, c in zip(begins, sizes, cutoffs)]
    for call in range(2):
        outs = (q1.Result * n)()
        note(f"case batch of {n} requests, call {call}")
        ctx.check(q1.lib.q1_run_batch(ctx.handle, reqs, n, None, outs), "q1_run_batch")
        for i in range(n):
            assert q1.result_to_list(outs[i]) == want[i], (call, i, begins[i], int(sizes[i]))
    t.free()


def test_batch_rows_per_lane_and_empty_requests(ctx):
    """The batched kernel cuts each request into blocks of 256 lanes x tuples_per_thread
    rows.  Small values give large requests many blocks; empty requests (first, between
    others and last) own no block, so the search for a block's request must pass over
    them.  Request sizes straddle the block size; every grid and rows per lane gives the
    results of the requests run alone.  Values that are not a multiple of 16, or above
    16384, are refused, and so are grids of 2^24 blocks or more (2^32 work-items)."""
    t = ctx.generate(q1.COMPRESSED, SEED + 33, 300_000)
    spans = [(0, 0), (5, 1), (0, 4095), (7, 4096), (9, 4097), (100, 0), (13, 65_549), (0, 300_000),
             (299_999, 1), (1234, 0), (50_001, 123_457), (3, 0)]
    cut = [729999 - 7 * i for i in range(len(spans))]
    reqs = [q1.Request(t.handle, b, c, cut[i], 0) for i, (b, c) in enumerate(spans)]
    want = [ctx.run(t, b, c, cut[i], q1.IN_REGISTERS) for i, (b, c) in enumerate(spans)]
    n = len(reqs)
    for tpt in (16, 48, 256, 0):
        for blocks in (1, 5, 0):
            outs = (q1.Result * n)()
            note(f"case batch rows per lane tpt={tpt} blocks={blocks}")
            ctx.check(q1.lib.q1_run_batch(ctx.handle, (q1.Request * n)(*reqs), n,
                                          C.byref(q1.make_launch(0, tpt, blocks)), outs), "q1_run_batch")
            for i in range(n):
                assert q1.result_to_list(outs[i]) == want[i], (tpt, blocks, i)
    for bad_tpt, bad_blocks in ((24, 0), (16400, 0), (0, 1 << 24), (0, (1 << 24) + 1)):
        outs = (q1.Result * n)()
        st = q1.lib.q1_run_batch(ctx.handle, (q1.Request * n)(*reqs), n,
                                 C.byref(q1.make_launch(0, bad_tpt, bad_blocks)), outs)
        assert st == q1.EINVAL, (bad_tpt, bad_blocks)
    t.free()


@pytest.mark.parametrize("layout", [q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED])
@pytest.mark.parametrize("spread", ["scattered", "everywhere"])
def test_enqueue_without_domain_flag_outside_the_domain(ctx, oracle, layout, spread):
    """q1_enqueue without the domain flag (the reference's launch contract) runs the
    any-value kernel that keeps out-of-domain sums in a per-lane private table, touched only
    from a lane's first out-of-domain row: rows outside the domain in a few scattered rows
    (most wavefront groups stay on the in-domain body) or in every fifth row (every lane
    uses its private table).  Every placement and several shapes, exact vs the oracle."""
    import torch
    rows = 200_003
    ref = q1.generate_host(q1.REFERENCE, SEED + 41, rows)
    rng = np.random.default_rng(41)
    bad = rng.choice(rows, 40, replace=False) if spread == "scattered" else np.arange(0, rows, 5)
    k = len(bad) // 4
    if layout == q1.COMPRESSED:
        ref["discount"][bad[:k]] = 200
        ref["tax"][bad[k:2 * k]] = 99
        ref["extended_price"][bad[2 * k:3 * k]] = 3_000_000_000
        ref["quantity"][bad[3 * k:]] = 9900
        dev = q1.compress_columns(ref, rows)
    elif layout == q1.UNCOMPRESSED:
        ref["discount"][bad[:k]] = -5
        ref["tax"][bad[k:2 * k]] = 100
        ref["extended_price"][bad[2 * k:3 * k]] = 1 << 45
        ref["quantity"][bad[3 * k:]] = 6301
        dev = {"ship_date": ref["ship_date"], "quantity": ref["quantity"].astype(np.int32),
               "extended_price": ref["extended_price"], "discount": ref["discount"].astype(np.int8),
               "tax": ref["tax"].astype(np.int8), "return_flag": ref["return_flag"],
               "line_status": ref["line_status"]}
    else:
        ref["discount"][bad[:k]] = -5
        ref["tax"][bad[k:2 * k]] = 1 << 40
        ref["extended_price"][bad[2 * k:3 * k]] = 1 << 45
        ref["quantity"][bad[3 * k:]] = 6301
        dev = ref
    want, _ = oracle.naive(ref, 0, rows, 729999)
    t = ctx.upload(layout, dev, rows)
    cols = t.device_columns()
    agg = torch.zeros(48, dtype=torch.int64, device="cuda")
    for p in range(5):
        for shape in (None, (64, 16), (256, 1024)):
            note(f"case enqueue any-value layout={q1.LAYOUT_NAMES[layout]} {spread} placement={p} shape={shape}")
            agg.zero_()
            torch.cuda.synchronize()
            l = q1.make_launch(*shape) if shape else None
            ctx.check(q1.lib.q1_enqueue(ctx.handle, p, layout, C.byref(l) if l else None, C.c_void_p(agg.data_ptr()),
                                        C.byref(cols), 0, rows, 729999, None), "q1_enqueue")
            ctx.synchronize()
            assert as_dicts(device_aggregates_to_result(agg)) == want, (p, shape)
    t.free()
