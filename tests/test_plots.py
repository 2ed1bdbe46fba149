"""The plots that replace scripts/plots.R, drawn from the committed autotuning table and a
synthetic bench line (CPU; skipped where matplotlib is absent), and the autotuner's
streaming lines."""
import json
import os

import pytest

from tpchq01_gpu_amd import PACKAGE_DIR, autotune, plots, q1

TABLE = os.path.join(PACKAGE_DIR, "tuning", "gfx950.json")
PNG = b"\x89PNG\r\n\x1a\n"


def _stream(chunk, streams, ms, rows=1_000_000):
    return dict(layout=q1.COMPRESSED, placement=q1.AUTO, rows=rows, chunk_rows=chunk, streams=streams, wall_ms=ms,
                kernel_ms=ms, gbps=rows * 9.375 / (ms * 1e-3) / 1e9)


BENCH = {"config": {"rows": 600_000_000}, "ms_per_step": 0.85,
         "cpu_baseline": {"value": 2.4e9, "cores": 16, "rows_6M": {"naive_1_thread": 1.5e8}},
         "host_resident": {"streaming": {"ms_per_query": 100.0},
                           "split": [{"cpu_fraction": 0.5, "wall_ms": 34.0}, {"cpu_fraction": 0.012, "wall_ms": 1.4}]},
         "placements": {"in_registers": {"kernel_ms": 1.02}, "shared_mem_per_thread": {"kernel_ms": 0.84}},
         "layouts": {"compressed_pushdown": {"kernel_ms": 0.67, "placement": "shared_mem_per_thread"},
                     "reference": {"kernel_ms": 3.4, "placement": "shared_mem_per_thread"}}}


def test_comparison_rows_and_options_table():
    rows = dict(plots.comparison_rows(BENCH))
    assert rows["CPU, 1 thread (rate at 6M rows)"] == pytest.approx(4.0)
    assert rows["CPU, 16 threads"] == pytest.approx(0.25)
    assert rows["GPU, columns in HBM"] == pytest.approx(0.85e-3)
    assert rows["GPU, columns streamed from host"] == pytest.approx(0.1)
    assert rows["CPU + GPU split (1.2% on CPU)"] == pytest.approx(1.4e-3)   # the faster split
    text = plots.options_table(BENCH)
    body = [l for l in text.splitlines() if l.startswith("| ") and "---" not in l][1:]
    ms = [float(l.split("|")[5]) for l in body]
    assert ms == sorted(ms, reverse=True) and len(body) == 6
    assert "| small (compressed) | yes | shared_mem_per_thread | no | 0.670 |" in text
    assert "| full (reference) | no | shared_mem_per_thread | no | 3.400 |" in text


def test_plots_render(tmp_path):
    pytest.importorskip("matplotlib")
    with open(TABLE) as f:
        table = json.load(f)
    plots.threads_heatmaps(table, str(tmp_path / "threads.png"))
    plots.comparison_bars(BENCH, str(tmp_path / "comparison.png"))
    assert not plots.streams_heatmap({"streaming_measurements": []}, str(tmp_path / "none.png"))
    table["streaming_measurements"] = [_stream(c, s, 10.0 + c / 1e6 + s) for c in (1 << 20, 1 << 21) for s in (1, 2, 4)]
    assert plots.streams_heatmap(table, str(tmp_path / "streams.png"))
    for name in ("threads.png", "comparison.png", "streams.png"):
        with open(tmp_path / name, "rb") as f:
            assert f.read(8) == PNG, name
    assert not (tmp_path / "none.png").exists()


def test_streaming_lines_and_report(tmp_path):
    with open(TABLE) as f:
        table = json.load(f)
    sres = [_stream(1 << 20, 1, 30.0), _stream(1 << 20, 4, 12.0), _stream(1 << 22, 4, 11.0), _stream(1 << 22, 8, 11.5)]
    autotune.write_tables(table["measurements"], str(tmp_path), table["device"], table["rows"],
                          table.get("batch_measurements", []), stream_results=sres)
    lines = [l.split() for l in open(tmp_path / "gfx950.txt") if not l.startswith("#")]
    assert ["streaming", str(q1.COMPRESSED), str(1 << 22), "4"] in lines
    with open(tmp_path / "gfx950.json") as f:
        written = json.load(f)
    assert written["streaming_measurements"] == sres
    text = autotune.report(written)
    assert "4,194,304 rows per chunk x 4 streams, 11.0 ms" in text
    assert "| 1,048,576 | 30.0 | 12.0 |  |" in text
    # merging keeps the old streaming sweep when a run does not sweep it again
    assert autotune.merge_results(sres, []) == sres
