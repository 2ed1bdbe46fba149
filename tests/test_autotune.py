"""The autotuner's tables: the committed tuning/gfx950.{txt,json} are exactly what
write_tables emits from the committed measurements (CPU), the report covers every tuned
(layout, placement), and a small sweep on the GPU writes a table that q1_init reads back
(scan shapes through q1_tuning_get, batch shapes through the batched launch)."""
import json
import os

import pytest

from tpchq01_gpu_amd import PACKAGE_DIR, autotune, q1

TUNING = os.path.join(PACKAGE_DIR, "tuning")


def test_committed_tables_are_write_tables_output(tmp_path):
    with open(os.path.join(TUNING, "gfx950.json")) as f:
        table = json.load(f)
    autotune.write_tables(table["measurements"], str(tmp_path), table["device"], table["rows"],
                          table.get("batch_measurements", []), stream_results=table.get("streaming_measurements", []))
    for name in ("gfx950.txt", "gfx950.json"):
        with open(os.path.join(TUNING, name)) as a, open(tmp_path / name) as b:
            assert a.read() == b.read(), name


def test_report_lists_every_tuned_shape():
    with open(os.path.join(TUNING, "gfx950.json")) as f:
        table = json.load(f)
    text = autotune.report(table)
    marks = 0
    for n in autotune.sizes_of(table["measurements"]):
        best, per_layout = autotune.best_table(table["measurements"], n)
        assert f"{n:,} rows per layout" in text
        for (l, p), r in best.items():
            assert f"| {q1.LAYOUT_NAMES[l]} | {q1.PLACEMENT_NAMES[p]}" in text
            assert f"| {r['tpb']} | {r['tpt']} | {r['kernel_ms']:.3f} |" in text
        marks += len(per_layout)
    assert text.count("**best**") == marks


def test_merge_keeps_what_was_not_swept():
    old = [dict(layout=0, placement=1, rows=10, kernel_ms=2.0), dict(layout=2, placement=1, rows=10, kernel_ms=3.0),
           dict(layout=2, placement=1, rows=99, kernel_ms=4.0)]
    new = [dict(layout=2, placement=1, rows=10, kernel_ms=1.0)]
    assert autotune.merge_results(old, new) == [old[0], old[2], new[0]]


def test_batch_shape_is_the_largest_grid_within_one_percent():
    def b(per_cu, tpt, ms, p=3):
        return dict(layout=2, placement=p, blocks_per_cu=per_cu, tpt=tpt, kernel_ms=ms)
    res = [b(3, 256, 1.103), b(4, 256, 1.111), b(8, 192, 1.104), b(8, 256, 1.110), b(2, 256, 1.226),
           b(6, 512, 1.30, p=0), b(8, 512, 1.40, p=0)]
    best = autotune.best_batch_table(res)
    assert (best[(2, 3)]["blocks_per_cu"], best[(2, 3)]["tpt"]) == (8, 192)
    assert (best[(2, 0)]["blocks_per_cu"], best[(2, 0)]["tpt"]) == (6, 512)   # 8 x 512 is 8 % slower
    assert autotune.best_batch_table(res, tie=0.0)[(2, 3)]["blocks_per_cu"] == 3


def _m(layout, placement, rows, tpb, tpt, ms):
    return dict(layout=layout, placement=placement, rows=rows, tpb=tpb, tpt=tpt, blocks=1, kernel_ms=ms,
                gbps=rows * q1.BITS_PER_ROW[layout] / 8 / (ms * 1e-3) / 1e9)


def test_several_row_counts_give_launch_at_lines(tmp_path):
    """Measurements at two row counts: the "launch" and "best" lines come from the larger,
    one "launch_at" line per row count and (layout, placement) carries that count's fastest
    shape, and the report has a table per row count."""
    big, small = 600_000_000, 6_000_000
    res = [_m(2, 3, big, 256, 128, 0.84), _m(2, 3, big, 64, 64, 0.90), _m(2, 0, big, 256, 1024, 1.02),
           _m(2, 3, small, 256, 128, 0.029), _m(2, 3, small, 64, 64, 0.025), _m(2, 0, small, 128, 32, 0.020)]
    best, per_layout = autotune.write_tables(res, str(tmp_path), "test", [big, small])
    assert per_layout[2]["placement"] == 3 and best[(2, 3)]["tpt"] == 128
    lines = [l.split() for l in open(tmp_path / "gfx950.txt") if not l.startswith("#")]
    assert ["launch", "2", "3", "256", "128"] in lines and ["launch", "2", "0", "256", "1024"] in lines
    assert ["best", "2", "3"] in lines                  # the small count's fastest placement is 0
    at = [l for l in lines if l[0] == "launch_at"]
    assert at == [["launch_at", str(small), "2", "0", "128", "32"], ["launch_at", str(small), "2", "3", "64", "64"],
                  ["launch_at", str(big), "2", "0", "256", "1024"], ["launch_at", str(big), "2", "3", "256", "128"]]
    with open(tmp_path / "gfx950.json") as f:
        table = json.load(f)
    assert table["rows"] == [big, small] and len(table["best"]) == 4
    text = autotune.report(table)
    assert f"{big:,} rows per layout" in text and f"{small:,} rows per layout" in text
    assert "| 64 | 64 | 0.025 |" in text and "| 256 | 128 | 0.840 |" in text
    assert text.count("**best**") == 2


@pytest.mark.gpu
def test_small_sweep_is_read_back(tmp_path, monkeypatch):
    from tpchq01_gpu_amd import workloads
    layout = q1.COMPRESSED
    placements = [q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM]
    with q1.Context(0) as ctx:
        cus = ctx.device_info().compute_units
        # two row counts at which every swept shape gives at least 4 blocks per CU (so a
        # launch of that many rows takes its shape unhalved)
        small = 4 * cus * 256 * 64 + 77
        big = 4 * small
        res = []
        for n in (big, small):
            res += autotune.sweep(ctx, layout, n, placements, 1, log=lambda s: None, threads=[64, 256], tuples=[32, 64])
        spec = [(s, n // 100, c) for s, n, c in workloads.batch_spec(8)]
        bres = autotune.sweep_batch(ctx, layout, spec, placements, 1,
                                    log=lambda s: None, blocks_per_cu=[1, 2],
                                    lane_rows={q1.IN_REGISTERS: [64, 128], q1.PER_THREAD_SHARED_MEM: [64, 128]})
    assert len(res) == 2 * 2 * 2 * 2 and len(bres) == 2 * 2 * 2
    best, per_layout = autotune.write_tables(res, str(tmp_path), "test", [big, small], bres)
    best_small, _ = autotune.best_table(res, small)
    bbest = autotune.best_batch_table(bres)
    monkeypatch.setenv("Q1_TUNING_FILE", str(tmp_path / "gfx950.txt"))
    with q1.Context(0) as ctx:
        for (l, p), r in best.items():
            got = ctx.tuning_get(l, p)
            assert (got.threads_per_block, got.tuples_per_thread) == (r["tpb"], r["tpt"]), (l, p)
            # a launch takes the shape tuned at the row count nearest its own, by ratio
            for n, want in ((big, r), (small, best_small[(l, p)]), (2 * small - 1, best_small[(l, p)]),
                            (2 * small + 1, r), (100 * big, r)):
                got = ctx.tuning_resolve(l, p, n)
                assert (got.threads_per_block, got.tuples_per_thread) == (want["tpb"], want["tpt"]), (l, p, n)
        # the batched kernel runs the layout's best placement with that placement's batch shape
        p = per_layout[layout]["placement"]
        want = bbest[(layout, p)]
        t = ctx.generate(layout, 0x51C00042, 100_000)
        reqs = (q1.Request * 2)(q1.Request(t.handle, 0, 100_000, 729999, 0), q1.Request(t.handle, 7, 5000, 729906, 0))
        outs = (q1.Result * 2)()
        ctx.check(q1.lib.q1_run_batch(ctx.handle, reqs, 2, None, outs), "q1_run_batch")
        tm = ctx.last_timing()
        assert (tm.placement, tm.num_blocks, tm.tuples_per_thread) == (p, want["blocks_per_cu"] * cus, want["tpt"])
        assert q1.result_to_list(outs[0]) == ctx.run(t, 0, 100_000, 729999, q1.IN_REGISTERS)
        assert q1.result_to_list(outs[1]) == ctx.run(t, 7, 5000, 729906, q1.IN_REGISTERS)
        t.free()


@pytest.mark.gpu
def test_resolve_takes_the_nearest_row_count(tmp_path, monkeypatch):
    """A hand-written table with shapes at 6M, 60M and 600M rows: q1_tuning_resolve takes
    the row count nearest by ratio, halves tuples_per_thread below 4 blocks per CU, skips
    malformed lines, and q1_tuning_set replaces the shape at every row count."""
    f = tmp_path / "t.txt"
    f.write_text("launch 2 3 256 128\n"
                 "launch_at 60000000 2 3 256 32\n"
                 "launch_at 6000000 2 3 64 64\n"
                 "launch_at 600000000 2 3 256 128\n"
                 "launch_at 0 2 3 128 16\n"          # no row count: ignored
                 "launch_at 1000 2 3 100 16\n"       # not a block size: ignored
                 "best 2 3\n")
    monkeypatch.setenv("Q1_TUNING_FILE", str(f))
    with q1.Context(0) as ctx:
        cus = ctx.device_info().compute_units

        def shape(n, p=q1.PER_THREAD_SHARED_MEM):
            l = ctx.tuning_resolve(q1.COMPRESSED, p, n)
            return l.threads_per_block, l.tuples_per_thread

        assert shape(6_000_000_000) == shape(600_000_000) == shape(200_000_000) == (256, 128)
        assert shape(180_000_000) == shape(60_000_000) == shape(19_000_000) == (256, 32)  # sqrt(6M x 60M) = 18.97M
        assert shape(18_000_000) == shape(6_000_000) == (64, 64)
        assert shape(6_000_000, q1.AUTO) == (64, 64)
        n = 4 * cus * 64 * 32                      # exactly 4 blocks per CU at 64 x 32
        assert shape(n) == (64, 32) and shape(n - 1) == (64, 16) and shape(2000) == (64, 16)
        got = ctx.tuning_get(q1.COMPRESSED, q1.PER_THREAD_SHARED_MEM)
        assert (got.threads_per_block, got.tuples_per_thread) == (256, 128)
        # placements without launch_at lines keep their one shape (halved for small ranges)
        other = ctx.tuning_get(q1.COMPRESSED, q1.IN_REGISTERS)
        assert shape(600_000_000, q1.IN_REGISTERS) == (other.threads_per_block, other.tuples_per_thread)
        ctx.tuning_set(q1.COMPRESSED, q1.PER_THREAD_SHARED_MEM, 128, 512)
        assert shape(600_000_000) == (128, 512) and shape(6_000_000)[0] == 128
        # a launch with no caller shape runs the resolved one
        t = ctx.generate(q1.COMPRESSED, 0x51C00043, 300_000)
        ctx.run(t, 0, 300_000, 729999, q1.PER_THREAD_SHARED_MEM)
        tm = ctx.last_timing()
        assert (tm.threads_per_block, tm.tuples_per_thread) == shape(300_000)
        t.free()


@pytest.mark.gpu
def test_streaming_sweep_is_read_back(tmp_path, monkeypatch, oracle):
    """A small sweep of the host-resident pipeline's rows per chunk x streams: its best pair
    becomes the table's streaming line, q1_run_streaming with chunk_rows = streams = 0 runs
    it (and matches the oracle), and a layout without a line keeps 16M rows x 4 streams."""
    rows, seed = 3_000_017, 0x51C00044
    with open(os.path.join(TUNING, "gfx950.json")) as f:
        table = json.load(f)
    with q1.Context(0) as ctx:
        sres = autotune.sweep_streaming(ctx, q1.COMPRESSED, rows, 1, seed=seed, log=lambda s: None,
                                        chunks=[1 << 19, 1 << 20], counts=[1, 3], host_threads=4)
    assert len(sres) == 4 and all(r["wall_ms"] > 0 for r in sres)
    autotune.write_tables(table["measurements"], str(tmp_path), table["device"], table["rows"],
                          table.get("batch_measurements", []), stream_results=sres)
    best = autotune.best_streaming(sres)[q1.COMPRESSED]
    monkeypatch.setenv("Q1_TUNING_FILE", str(tmp_path / "gfx950.txt"))
    with q1.Context(0) as ctx:
        assert ctx.tuning_get_streaming(q1.COMPRESSED) == (best["chunk_rows"], best["streams"])
        assert ctx.tuning_get_streaming(q1.UNCOMPRESSED) == (1 << 24, 4)
        want, _ = oracle.naive(q1.generate_host(q1.REFERENCE, seed, rows), 0, rows, 729906)
        got = ctx.run_streaming(q1.COMPRESSED, q1.generate_host(q1.COMPRESSED, seed, rows), rows, 729906,
                                q1.AUTO, q1.make_launch(flags=q1.LAUNCH_ASSUME_TPCH_DOMAIN), 0, 0)
        assert [dict(sum_qty=g.sum_qty, sum_base_price=g.sum_base_price, sum_disc_price=g.sum_disc_price,
                     sum_charge=g.sum_charge, sum_disc=g.sum_disc, count=g.count) for g in got] == want
