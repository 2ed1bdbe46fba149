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
                          table.get("batch_measurements", []))
    for name in ("gfx950.txt", "gfx950.json"):
        with open(os.path.join(TUNING, name)) as a, open(tmp_path / name) as b:
            assert a.read() == b.read(), name


def test_report_lists_every_tuned_shape():
    with open(os.path.join(TUNING, "gfx950.json")) as f:
        table = json.load(f)
    text = autotune.report(table)
    best, per_layout = autotune.best_table(table["measurements"])
    for (l, p), r in best.items():
        assert f"| {q1.LAYOUT_NAMES[l]} | {q1.PLACEMENT_NAMES[p]}" in text
        assert f"| {r['tpb']} | {r['tpt']} | {r['kernel_ms']:.3f} |" in text
    assert text.count("**best**") == len(per_layout)


def test_merge_keeps_what_was_not_swept():
    old = [dict(layout=0, placement=1, kernel_ms=2.0), dict(layout=2, placement=1, kernel_ms=3.0)]
    new = [dict(layout=2, placement=1, kernel_ms=1.0)]
    assert autotune.merge_results(old, new) == [old[0], new[0]]


@pytest.mark.gpu
def test_small_sweep_is_read_back(tmp_path, monkeypatch):
    from tpchq01_gpu_amd import workloads
    layout = q1.COMPRESSED
    with q1.Context(0) as ctx:
        cus = ctx.device_info().compute_units
        rows = 2 * cus * 256 * 64 + 77      # every swept shape leaves at least 2 blocks per CU
        res = autotune.sweep(ctx, layout, rows, [q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM], 1, log=lambda s: None,
                             threads=[64, 256], tuples=[32, 64])
        spec = [(s, n // 100, c) for s, n, c in workloads.batch_spec(8)]
        bres = autotune.sweep_batch(ctx, layout, spec, [q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM], 1,
                                    log=lambda s: None, blocks_per_cu=[1, 2],
                                    lane_rows={q1.IN_REGISTERS: [64, 128], q1.PER_THREAD_SHARED_MEM: [64, 128]})
    assert len(res) == 2 * 2 * 2 and len(bres) == 2 * 2 * 2
    best, per_layout = autotune.write_tables(res, str(tmp_path), "test", rows, bres)
    bbest, _ = autotune.best_table(bres)
    monkeypatch.setenv("Q1_TUNING_FILE", str(tmp_path / "gfx950.txt"))
    with q1.Context(0) as ctx:
        for (l, p), r in best.items():
            got = ctx.tuning_get(l, p)
            assert (got.threads_per_block, got.tuples_per_thread) == (r["tpb"], r["tpt"]), (l, p)
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
