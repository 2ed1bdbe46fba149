"""Plots of the autotuner's sweeps and of a bench line (replaces scripts/plots.R).

plots.R draws, from the reference's result CSVs:
  - threads.pdf:    a heatmap of kernel time over threads per block x values per thread
                    (scripts/plots.R:58-86);
  - streams.pdf:    a heatmap of time over tuples per kernel launch x streams (:32-56);
  - comparison.pdf: bars of the Q1 time on one CPU thread, one CPU socket, the GPU and
                    CPU + GPU (:88-104);
  - a LaTeX table of the time per option combination (coprocessing, filter pushdown,
                    table placement, compressed types) (:106-121).
Here the same views come from tuning/<arch>.json (the autotuner's measurements) and from a
bench.py JSON line:
  - threads.png:    per layout (rows) and placement (columns) a heatmap of the fraction of
                    8 TB/s over threads per block x tuples per thread, at one row count;
  - streams.png:    the host-resident pipeline's ms per query over rows per chunk x streams;
  - comparison.png: time for the bench's 600M rows on one CPU thread, on all host cores,
                    on the GPU with device-resident columns, streamed from host memory, and
                    split between CPU and GPU;
  - options.md:     the bench's time per (compressed types, filter pushdown, placement,
                    coprocessing) combination.

Usage: python -m tpchq01_gpu_amd.plots [--table tuning/gfx950.json] [--bench BENCH.json]
       [--rows N] [--out tuning/plots]
"""
from __future__ import annotations

import argparse
import json
import math
import os

from . import PACKAGE_DIR, q1

HBM_PEAK_GBS = 8000.0


def _pyplot():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    return plt


def threads_heatmaps(table: dict, out_png: str, rows: int | None = None) -> None:
    """One heatmap per (layout, placement) swept at `rows` rows (default: the largest row
    count in the table): threads per block x tuples per thread, coloured by the fraction of
    8 TB/s; the best shape of each panel is outlined."""
    plt = _pyplot()
    meas = table["measurements"]
    if rows is None:
        rows = max(r["rows"] for r in meas)
    meas = [r for r in meas if r["rows"] == rows]
    layouts = sorted({r["layout"] for r in meas})
    placements = sorted({r["placement"] for r in meas})
    tpbs = sorted({r["tpb"] for r in meas})
    tpts = sorted({r["tpt"] for r in meas})
    fig, axes = plt.subplots(len(layouts), len(placements), figsize=(3.2 * len(placements), 2.6 * len(layouts) + 1.0),
                             squeeze=False, sharex=True, sharey=True, layout="constrained")
    image = None
    for i, l in enumerate(layouts):
        for j, p in enumerate(placements):
            ax = axes[i][j]
            grid = [[math.nan] * len(tpts) for _ in tpbs]
            best = None
            for r in meas:
                if r["layout"] == l and r["placement"] == p:
                    a, b = tpbs.index(r["tpb"]), tpts.index(r["tpt"])
                    grid[a][b] = r["gbps"] / HBM_PEAK_GBS
                    if best is None or r["kernel_ms"] < best["kernel_ms"]:
                        best = r
            image = ax.imshow(grid, origin="lower", aspect="auto", vmin=0.0, vmax=1.0, cmap="viridis")
            if best is not None:
                ax.add_patch(plt.Rectangle((tpts.index(best["tpt"]) - 0.5, tpbs.index(best["tpb"]) - 0.5), 1, 1,
                                           fill=False, edgecolor="red", linewidth=2))
                ax.set_title(f"{q1.LAYOUT_NAMES[l]}\n{q1.PLACEMENT_NAMES[p]}: {best['gbps'] / HBM_PEAK_GBS:.0%}",
                             fontsize=8)
            ax.set_xticks(range(len(tpts)))
            ax.set_xticklabels([str(t) for t in tpts], rotation=90, fontsize=6)
            ax.set_yticks(range(len(tpbs)))
            ax.set_yticklabels([str(t) for t in tpbs], fontsize=7)
            if i == len(layouts) - 1:
                ax.set_xlabel("tuples per thread", fontsize=7)
            if j == 0:
                ax.set_ylabel("threads per block", fontsize=7)
    fig.suptitle(f"Q1 scan, fraction of 8 TB/s HBM peak, {rows:,} device-resident rows ({table.get('device', '')})",
                 fontsize=10)
    fig.colorbar(image, ax=axes, shrink=0.6, label="fraction of 8 TB/s")
    fig.savefig(out_png, dpi=80)
    plt.close(fig)


def streams_heatmap(table: dict, out_png: str, layout: int = q1.COMPRESSED) -> bool:
    """The host-resident pipeline's end-to-end ms per query over rows per chunk x streams
    (False when the table has no streaming sweep for `layout`)."""
    meas = [r for r in table.get("streaming_measurements", []) if r["layout"] == layout]
    if not meas:
        return False
    plt = _pyplot()
    chunks = sorted({r["chunk_rows"] for r in meas})
    counts = sorted({r["streams"] for r in meas})
    grid = [[math.nan] * len(chunks) for _ in counts]
    for r in meas:
        grid[counts.index(r["streams"])][chunks.index(r["chunk_rows"])] = r["wall_ms"]
    fig, ax = plt.subplots(figsize=(7, 4.5))
    image = ax.imshow(grid, origin="lower", aspect="auto", cmap="viridis_r")
    values = [v for row in grid for v in row if not math.isnan(v)]
    mid = (min(values) + max(values)) / 2
    for a in range(len(counts)):
        for b in range(len(chunks)):
            if not math.isnan(grid[a][b]):
                ax.text(b, a, f"{grid[a][b]:.0f}", ha="center", va="center", fontsize=8,
                        color="white" if grid[a][b] > mid else "black")
    ax.set_xticks(range(len(chunks)))
    ax.set_xticklabels([f"{c >> 20}M" for c in chunks])
    ax.set_yticks(range(len(counts)))
    ax.set_yticklabels([str(c) for c in counts])
    ax.set_xlabel("rows per chunk (per kernel launch)")
    ax.set_ylabel("streams")
    ax.set_title(f"{q1.LAYOUT_NAMES[layout]} columns in pinned host memory, {meas[0]['rows']:,} rows:\n"
                 f"ms per query (copies and kernels overlapped)", fontsize=10)
    fig.colorbar(image, ax=ax, label="ms per query")
    fig.tight_layout()
    fig.savefig(out_png, dpi=80)
    plt.close(fig)
    return True


def comparison_rows(bench: dict) -> list[tuple[str, float]]:
    """(label, seconds) for the bench's headline rows: CPU one thread and all threads (the
    oracle's KernelNaive restatement in the reference layout), the GPU on device-resident
    columns, the host-resident streamed pipeline and the CPU/GPU split."""
    rows = bench["config"]["rows"]
    cpu = bench.get("cpu_baseline", {})
    out = []
    one = cpu.get("rows_6M", {}).get("naive_1_thread")
    if one:
        out.append(("CPU, 1 thread (rate at 6M rows)", rows / one))
    if cpu.get("value"):
        out.append((f"CPU, {cpu.get('cores', '?')} threads", rows / cpu["value"]))
    out.append(("GPU, columns in HBM", bench["ms_per_step"] * 1e-3))
    host = bench.get("host_resident", {})
    if host.get("streaming"):
        out.append(("GPU, columns streamed from host", host["streaming"]["ms_per_query"] * 1e-3))
    splits = host.get("split", [])
    if splits:
        s = min(splits, key=lambda x: x["wall_ms"])
        out.append((f"CPU + GPU split ({s['cpu_fraction']:.1%} on CPU)", s["wall_ms"] * 1e-3))
    return out


def comparison_bars(bench: dict, out_png: str) -> None:
    plt = _pyplot()
    data = comparison_rows(bench)
    fig, ax = plt.subplots(figsize=(8, 3.6))
    labels = [d[0] for d in data]
    secs = [d[1] for d in data]
    ax.barh(range(len(data)), secs, color="#777777")
    ax.set_xscale("log")
    ax.set_yticks(range(len(data)))
    ax.set_yticklabels(labels)
    ax.invert_yaxis()
    for i, v in enumerate(secs):
        ax.text(v, i, f"  {v * 1e3:.2f} ms" if v < 1 else f"  {v:.2f} s", va="center", fontsize=8)
    ax.set_xlabel("seconds for the query (log scale)")
    ax.set_title(f"TPC-H Q1 over {bench['config']['rows']:,} rows")
    ax.set_xlim(min(secs) / 3, max(secs) * 20)
    fig.tight_layout()
    fig.savefig(out_png, dpi=80)
    plt.close(fig)


def options_table(bench: dict) -> str:
    """Markdown: the bench's device time per (compressed types, filter pushdown,
    placement, coprocessing) combination, slowest first, as plots.R's option table."""
    rows = []
    for p, r in bench.get("placements", {}).items():
        rows.append(("small (compressed)", "no", p, "no", r["kernel_ms"]))
    lay = bench.get("layouts", {})
    for key, types, push in (("compressed_pushdown", "small (compressed)", "yes"),
                             ("uncompressed", "full (uncompressed)", "no"),
                             ("reference", "full (reference)", "no")):
        if key in lay:
            rows.append((types, push, lay[key].get("placement", "shared_mem_per_thread"), "no", lay[key]["kernel_ms"]))
    for s in bench.get("host_resident", {}).get("split", []):
        rows.append(("small (compressed)", "no", "auto", f"yes ({s['cpu_fraction']:.1%})", s["wall_ms"]))
    rows.sort(key=lambda r: -r[4])
    n = bench["config"]["rows"]
    lines = [f"Time per option combination, {n:,} rows (device ms; coprocessing rows: wall ms of the split)", "",
             "| types | filter pushdown | placement | coprocessing | ms |", "|---|---|---|---|---:|"]
    lines += [f"| {t} | {f} | {p} | {c} | {ms:.3f} |" for t, f, p, c, ms in rows]
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--table", default=os.path.join(PACKAGE_DIR, "tuning", "gfx950.json"))
    ap.add_argument("--bench", help="a file whose last line is a bench.py JSON line")
    ap.add_argument("--rows", type=int, help="row count of the threads heatmaps (default: the largest swept)")
    ap.add_argument("--out", default=os.path.join(PACKAGE_DIR, "tuning", "plots"))
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    with open(args.table) as f:
        table = json.load(f)
    written = [os.path.join(args.out, "threads.png")]
    threads_heatmaps(table, written[0], args.rows)
    if streams_heatmap(table, os.path.join(args.out, "streams.png")):
        written.append(os.path.join(args.out, "streams.png"))
    if args.bench:
        with open(args.bench) as f:
            bench = json.loads(f.read().strip().splitlines()[-1])
        written.append(os.path.join(args.out, "comparison.png"))
        comparison_bars(bench, written[-1])
        written.append(os.path.join(args.out, "options.md"))
        with open(written[-1], "w") as f:
            f.write(options_table(bench))
    for w in written:
        print(w)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
