"""Launch-shape autotuner for the Q1 kernels (replaces the reference's
scripts/try_grid_params.py and the hand-entered optimum table of
scripts/benchmark_with_optimal_params.py:39-55).

Scan shapes: for every row count asked (by default 600M, 60M and 6M rows) and every
(layout, placement) it sweeps threads_per_block x tuples_per_thread, times the kernel alone
with HIP events (q1_last_timing().kernel_ms, the minimum over repetitions), skips shapes
that leave fewer than 2 blocks per compute unit (as try_grid_params.py:78-80 does), and
rotates between table copies when the table is smaller than the 256 MB Infinity Cache so
the timing reflects HBM.  The best shape differs with the row count: shorter launches
favour fewer rows per lane, whose last round of blocks leaves less of the chip idle.

Batch shapes: for each batch placement (the register table and the LDS table) it sweeps
the batched kernel's grid (blocks per compute unit) x rows per lane on the 256-request
batch workload (uneven log-uniform requests, their own cutoffs), timing q1_run_batch's
kernel; of the shapes within 1 % of the fastest it keeps the one with the most blocks per
compute unit.

Streaming (--streaming): the host-resident pipeline's rows per chunk x streams over 600M
compressed rows in pinned host memory, end-to-end time per query (the reference's
--tuples-per-kernel-launch and --streams).

It writes
  tuning/<arch>.txt    "launch <layout> <placement> <tpb> <tpt>" and "best <layout> <placement>"
                       lines (the largest row count swept), "launch_at <rows> <layout>
                       <placement> <tpb> <tpt>" lines (every row count, when several were
                       swept; a launch takes the one nearest its own row count) and
                       "batch <layout> <placement> <blocks per CU> <rows per lane>"
                       and "streaming <layout> <chunk rows> <streams>" lines, read by
                       q1_init (or Q1_TUNING_FILE)
  tuning/<arch>.json   every measurement
and `--report` prints the optimum per row count, layout and placement as tables
(tuning/<arch>_report.md), the place of the reference's per-configuration optimum list
(scripts/benchmark_with_optimal_params.py:39-55); tpchq01_gpu_amd.plots draws the sweeps
(scripts/plots.R).

Usage: python -m tpchq01_gpu_amd.autotune [--rows N[,N...]] [--layouts compressed,...] [--batch]
       [--streaming] [--merge] [--report]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

from . import PACKAGE_DIR, q1

THREADS = [64, 128, 256]
TUPLES = [16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]
ROWS = [600_000_000, 60_000_000, 6_000_000]
BATCH_BLOCKS_PER_CU = [1, 2, 3, 4, 6, 8]
STREAM_CHUNKS = [1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26]
STREAM_COUNTS = [1, 2, 3, 4, 6, 8]
BATCH_LANE_ROWS = {q1.IN_REGISTERS: [256, 512, 1024, 2048], q1.PER_THREAD_SHARED_MEM: [64, 128, 192, 256]}
MALL_BYTES = 256 << 20
HBM_PEAK_GBS = 8000.0
LAYOUT_KEYS = {"reference": q1.REFERENCE, "uncompressed": q1.UNCOMPRESSED, "compressed": q1.COMPRESSED,
               "compressed_pushdown": q1.COMPRESSED_PUSHDOWN}


def sweep(ctx: q1.Context, layout: int, rows: int, placements, reps: int, seed: int = 0x51C00010,
          log=print, threads=THREADS, tuples=TUPLES):
    algo = rows * q1.BITS_PER_ROW[layout] // 8
    copies = 1 if algo >= 4 * MALL_BYTES else max(2, (4 * MALL_BYTES) // max(algo, 1) + 1)
    copies = min(copies, 8)
    if layout == q1.COMPRESSED_PUSHDOWN:
        # the pushdown bitmap is built on the device from a compressed table (which the
        # pushdown table keeps alive)
        tables = [ctx.build_filter(ctx.generate(q1.COMPRESSED, seed + i, rows), 729999) for i in range(copies)]
    else:
        tables = [ctx.generate(layout, seed + i, rows) for i in range(copies)]
    cus = ctx.device_info().compute_units
    results = []
    try:
        for p in placements:
            for tpb in threads:
                for tpt in tuples:
                    blocks = -(-rows // (tpb * tpt))
                    if blocks < 2 * cus:
                        continue
                    launch = q1.make_launch(tpb, tpt)
                    best = float("inf")
                    for r in range(reps):
                        ctx.run_raw(tables[r % copies], 0, rows, 729999, p, launch)
                        best = min(best, ctx.last_timing().kernel_ms)
                    gbs = algo / (best * 1e-3) / 1e9
                    results.append(dict(layout=layout, placement=p, tpb=tpb, tpt=tpt, blocks=blocks,
                                        kernel_ms=best, gbps=gbs, rows=rows))
                    log(f"{q1.LAYOUT_NAMES[layout]:>20} {q1.PLACEMENT_NAMES[p]:>24} tpb={tpb:4d} tpt={tpt:5d} "
                        f"blocks={blocks:6d} {best:8.3f} ms {gbs:7.1f} GB/s")
    finally:
        for t in tables:
            t.free()
    return results


def sweep_batch(ctx: q1.Context, layout: int, spec, placements, reps: int, log=print,
                blocks_per_cu=BATCH_BLOCKS_PER_CU, lane_rows=BATCH_LANE_ROWS):
    """The batched kernel's grid x rows per lane on the requests of `spec` ((seed, rows,
    cutoff) per request, each on its own table), per batch placement."""
    tables = [ctx.generate(layout, s, n) for s, n, _ in spec]
    n = len(spec)
    reqs = (q1.Request * n)(*[q1.Request(tables[r].handle, 0, spec[r][1], spec[r][2], 0) for r in range(n)])
    outs = (q1.Result * n)()
    rows = sum(s[1] for s in spec)
    algo = rows * q1.BITS_PER_ROW[layout] // 8
    cus = ctx.device_info().compute_units
    results = []
    try:
        for p in placements:
            ctx.set_best_placement(layout, p)    # the batch runs the layout's best placement
            for per_cu in blocks_per_cu:
                for lr in lane_rows[p]:
                    launch = q1.make_launch(0, lr, per_cu * cus)
                    best = float("inf")
                    for r in range(reps + 1):    # the first call captures the graph
                        ctx.check(q1.lib.q1_run_batch(ctx.handle, reqs, n, C.byref(launch), outs), "q1_run_batch")
                        if r:
                            best = min(best, ctx.last_timing().kernel_ms)
                    gbs = algo / (best * 1e-3) / 1e9
                    results.append(dict(layout=layout, placement=p, blocks_per_cu=per_cu, tpt=lr,
                                        blocks=per_cu * cus, requests=n, rows=rows, kernel_ms=best, gbps=gbs))
                    log(f"batch {q1.LAYOUT_NAMES[layout]:>20} {q1.PLACEMENT_NAMES[p]:>24} blocks/CU={per_cu} "
                        f"tpt={lr:5d} {best:8.3f} ms {gbs:7.1f} GB/s")
    finally:
        for t in tables:
            t.free()
    return results


def sizes_of(results):
    """The row counts measured, largest first."""
    return sorted({r["rows"] for r in results}, reverse=True)


def sweep_streaming(ctx: q1.Context, layout: int, rows: int, reps: int, seed: int = 0x51C00011, log=print,
                    chunks=STREAM_CHUNKS, counts=STREAM_COUNTS, host_threads: int = 16):
    """The host-resident pipeline (q1_run_streaming): rows per chunk x streams, on `rows`
    rows generated into pinned host memory, end-to-end wall time per query (the minimum
    over repetitions after one untimed call that sets up the staging buffers); the
    reference sweeps the same two knobs (--tuples-per-kernel-launch, --streams,
    scripts/try_grid_params.py:37-38, scripts/plots.R's streams plot)."""
    algo = rows * q1.BITS_PER_ROW[layout] // 8
    cols = ctx.pinned_columns(layout, rows)
    results = []
    try:
        q1.fill_host(layout, seed, cols, rows, 0, host_threads)
        dom = q1.make_launch(flags=q1.LAUNCH_ASSUME_TPCH_DOMAIN)   # generated rows lie in the TPC-H domain
        for chunk in chunks:
            for streams in counts:
                if chunk > rows:
                    continue
                best = float("inf")
                for r in range(reps + 1):
                    t0 = time.perf_counter()
                    ctx.run_streaming(layout, cols, rows, 729999, q1.AUTO, dom, chunk, streams)
                    if r:
                        best = min(best, (time.perf_counter() - t0) * 1e3)
                gbs = algo / (best * 1e-3) / 1e9
                results.append(dict(layout=layout, placement=q1.AUTO, rows=rows, chunk_rows=chunk, streams=streams,
                                    wall_ms=best, kernel_ms=best, gbps=gbs))
                log(f"streaming {q1.LAYOUT_NAMES[layout]:>20} chunk={chunk:9d} streams={streams} {best:8.2f} ms "
                    f"{gbs:6.1f} GB/s")
    finally:
        ctx.free_pinned_columns(cols)
    return results


def best_table(results, rows=None):
    """The fastest shape per (layout, placement) at `rows` rows (default: the largest row
    count measured), and the fastest placement per layout among those."""
    if rows is None and results:
        rows = sizes_of(results)[0]
    best = {}
    for r in results:
        if r["rows"] != rows:
            continue
        k = (r["layout"], r["placement"])
        if k not in best or r["kernel_ms"] < best[k]["kernel_ms"]:
            best[k] = r
    per_layout = {}
    for (l, p), r in best.items():
        if l not in per_layout or r["kernel_ms"] < per_layout[l]["kernel_ms"]:
            per_layout[l] = r
    return best, per_layout


def merge_results(old, new):
    """Measurements of `old` whose (layout, placement, rows) `new` does not sweep again, plus
    `new`."""
    swept = {(r["layout"], r["placement"], r["rows"]) for r in new}
    return [r for r in old if (r["layout"], r["placement"], r["rows"]) not in swept] + list(new)


BATCH_TIE = 0.01


def best_batch_table(batch_results, tie=BATCH_TIE):
    """The batched kernel's shape per (layout, placement): of the shapes within `tie` of the
    fastest, the one with the most blocks per compute unit (then the fastest of those).
    Grids that leave wave slots empty measured within 1 % of full ones on the batch alone,
    but 5 % slower with other tables resident too (1.18 against 1.125 ms for 3 against 8
    blocks per CU), so among shapes the sweep cannot tell apart the full grid is kept."""
    fastest = {}
    for r in batch_results:
        k = (r["layout"], r["placement"])
        fastest[k] = min(fastest.get(k, float("inf")), r["kernel_ms"])
    best = {}
    for r in batch_results:
        k = (r["layout"], r["placement"])
        if r["kernel_ms"] > fastest[k] * (1 + tie):
            continue
        if k not in best or (r["blocks_per_cu"], -r["kernel_ms"]) > (best[k]["blocks_per_cu"], -best[k]["kernel_ms"]):
            best[k] = r
    return best


def best_streaming(stream_results):
    """The fastest (chunk rows, streams) per layout."""
    best = {}
    for r in stream_results:
        if r["layout"] not in best or r["wall_ms"] < best[r["layout"]]["wall_ms"]:
            best[r["layout"]] = r
    return best


def write_tables(results, out_dir: str, device: str, rows, batch_results=(), arch: str = "gfx950",
                 stream_results=()):
    """Writes tuning/<arch>.txt (read by q1_init) and tuning/<arch>.json.  `rows` (one row
    count or a list) is recorded as the sweep's; the lines come from the measurements."""
    os.makedirs(out_dir, exist_ok=True)
    sizes = sizes_of(results)
    best, per_layout = best_table(results)
    bbest = best_batch_table(batch_results)
    sbest = best_streaming(stream_results)
    with open(os.path.join(out_dir, f"{arch}.txt"), "w") as f:
        swept = ", ".join(str(n) for n in sizes)
        f.write(f"# written by tpchq01_gpu_amd.autotune on {device}, every layout swept at {swept} "
                f"device-resident rows:\n")
        f.write("#   launch <layout> <placement> <threads_per_block> <tuples_per_thread>   /   best <layout> <placement>\n")
        if len(sizes) > 1:
            f.write("#   launch_at <rows> <layout> <placement> <threads_per_block> <tuples_per_thread>"
                    "   (a launch takes the row count nearest its own)\n")
        f.write("#   batch <layout> <placement> <blocks_per_cu> <tuples_per_thread>   (the batched kernel, 256 threads)\n")
        if sbest:
            f.write("#   streaming <layout> <chunk_rows> <streams>   (host-resident columns, q1_run_streaming)\n")
        f.write("# (layouts 0 reference, 1 uncompressed, 2 compressed, 3 compressed + pushdown bitmap)\n")
        for (l, p), r in sorted(best.items()):
            f.write(f"launch {l} {p} {r['tpb']} {r['tpt']}\n")
        if len(sizes) > 1:
            for n in sorted(sizes):
                for (l, p), r in sorted(best_table(results, n)[0].items()):
                    f.write(f"launch_at {n} {l} {p} {r['tpb']} {r['tpt']}\n")
        for l, r in sorted(per_layout.items()):
            f.write(f"best {l} {r['placement']}\n")
        for (l, p), r in sorted(bbest.items()):
            f.write(f"batch {l} {p} {r['blocks_per_cu']} {r['tpt']}\n")
        for l, r in sorted(sbest.items()):
            f.write(f"streaming {l} {r['chunk_rows']} {r['streams']}\n")
    with open(os.path.join(out_dir, f"{arch}.json"), "w") as f:
        extra = {"streaming_measurements": list(stream_results)} if stream_results else {}
        json.dump({"device": device, "rows": rows, "measurements": list(results),
                   "batch_measurements": list(batch_results), **extra,
                   "best": [{**r, "layout": q1.LAYOUT_NAMES[l], "placement": q1.PLACEMENT_NAMES[p]}
                            for n in sizes for (l, p), r in sorted(best_table(results, n)[0].items())],
                   "batch_best": [{**r, "layout": q1.LAYOUT_NAMES[l], "placement": q1.PLACEMENT_NAMES[p]}
                                  for (l, p), r in sorted(bbest.items())]}, f, indent=1)
    return best, per_layout


def report(table: dict) -> str:
    """Markdown: per row count swept, the optimum launch per layout and placement, with its
    kernel time, rate and fraction of HBM peak and a bar per placement; then the batched
    kernel's optimum."""
    sizes = sizes_of(table["measurements"])
    lines = [f"# Tuned launch shapes on {table.get('device', '?')}"]
    for n in sizes:
        best, per_layout = best_table(table["measurements"], n)
        lines += ["", f"{n:,} rows per layout:", "",
                  "| layout | placement | threads/block | tuples/thread | kernel ms | GB/s | % of 8 TB/s | |",
                  "|---|---|---|---|---|---|---|---|"]
        for (l, p), r in sorted(best.items()):
            frac = r["gbps"] / HBM_PEAK_GBS
            mark = " **best**" if per_layout[l] is r else ""
            lines.append(f"| {q1.LAYOUT_NAMES[l]} | {q1.PLACEMENT_NAMES[p]}{mark} | {r['tpb']} | {r['tpt']} | "
                         f"{r['kernel_ms']:.3f} | {r['gbps']:.0f} | {100 * frac:.1f} | `{'#' * round(40 * frac)}` |")
    bbest = best_batch_table(table.get("batch_measurements", []))
    if bbest:
        lines += ["", "Batched kernel (256 threads per block; of the shapes within 1 % of the fastest, the largest grid):", "",
                  "| layout | placement | blocks/CU | tuples/thread | requests | rows | kernel ms | GB/s | % of 8 TB/s |",
                  "|---|---|---|---|---|---|---|---|---|"]
        for (l, p), r in sorted(bbest.items()):
            lines.append(f"| {q1.LAYOUT_NAMES[l]} | {q1.PLACEMENT_NAMES[p]} | {r['blocks_per_cu']} | {r['tpt']} | "
                         f"{r['requests']} | {r['rows']} | {r['kernel_ms']:.3f} | {r['gbps']:.0f} | "
                         f"{100 * r['gbps'] / HBM_PEAK_GBS:.1f} |")
    sres = table.get("streaming_measurements", [])
    if sres:
        lines += ["", "Host-resident columns streamed to the device (`q1_run_streaming`, end-to-end ms per query):"]
        for l, b in sorted(best_streaming(sres).items()):
            mine = [r for r in sres if r["layout"] == l]
            chunks = sorted({r["chunk_rows"] for r in mine})
            counts = sorted({r["streams"] for r in mine})
            ms = {(r["chunk_rows"], r["streams"]): r["wall_ms"] for r in mine}
            lines += ["", f"{q1.LAYOUT_NAMES[l]}, {b['rows']:,} rows; best {b['chunk_rows']:,} rows per chunk x "
                          f"{b['streams']} streams, {b['wall_ms']:.1f} ms ({b['gbps']:.1f} GB/s):", "",
                      "| rows per chunk | " + " | ".join(f"{c} streams" for c in counts) + " |",
                      "|---|" + "---|" * len(counts)]
            for ch in chunks:
                cells = [f"{ms[(ch, c)]:.1f}" if (ch, c) in ms else "" for c in counts]
                lines.append(f"| {ch:,} | " + " | ".join(cells) + " |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", default=",".join(str(n) for n in ROWS),
                    help="row counts to sweep the scan shapes at (comma-separated)")
    ap.add_argument("--layouts", default="compressed", help="layouts whose scan shapes are swept ('' = none)")
    ap.add_argument("--placements", default="in_registers,in_registers_per_thread,local_mem,shared_mem_per_thread,global")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--batch", action="store_true", help="also sweep the batched kernel's shapes (compressed)")
    ap.add_argument("--batch-requests", type=int, default=256)
    ap.add_argument("--streaming", action="store_true",
                    help="also sweep the host-resident pipeline's rows per chunk x streams (compressed)")
    ap.add_argument("--streaming-rows", type=int, default=600_000_000)
    ap.add_argument("--merge", action="store_true",
                    help="keep the existing table's measurements of what this run does not sweep")
    ap.add_argument("--report", action="store_true", help="print (and write) the report of the existing table")
    ap.add_argument("--out", default=os.path.join(PACKAGE_DIR, "tuning"))
    args = ap.parse_args(argv)
    json_path = os.path.join(args.out, "gfx950.json")
    if args.report:
        with open(json_path) as f:
            text = report(json.load(f))
        with open(os.path.join(args.out, "gfx950_report.md"), "w") as f:
            f.write(text)
        print(text, end="")
        return 0
    layouts = [LAYOUT_KEYS[x] for x in args.layouts.split(",") if x]
    placements = [q1.PLACEMENTS[x] for x in args.placements.split(",")]
    old = {"measurements": [], "batch_measurements": [], "streaming_measurements": []}
    if args.merge and os.path.exists(json_path):
        with open(json_path) as f:
            old = json.load(f)
    t0 = time.time()
    out = lambda s: print(s, flush=True)
    with q1.Context(0) as ctx:
        info = ctx.device_info()
        device = f"MI355X ({info.arch.decode().split(':')[0]})"
        results = []
        rows = [int(x) for x in args.rows.split(",") if x]
        for n in rows:
            for l in layouts:
                results += sweep(ctx, l, n, placements, args.reps, log=out)
        batch_results = []
        if args.batch:
            from . import workloads
            batch_results = sweep_batch(ctx, q1.COMPRESSED, workloads.batch_spec(args.batch_requests),
                                        [q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM], args.reps, log=out)
        stream_results = []
        if args.streaming:
            stream_results = sweep_streaming(ctx, q1.COMPRESSED, args.streaming_rows, args.reps, log=out)
    results = merge_results(old.get("measurements", []), results)
    batch_results = merge_results(old.get("batch_measurements", []), batch_results)
    stream_results = merge_results(old.get("streaming_measurements", []), stream_results)
    best, per_layout = write_tables(results, args.out, device, rows if len(rows) > 1 else rows[0], batch_results,
                                    stream_results=stream_results)
    for l, r in per_layout.items():
        print(f"best {q1.LAYOUT_NAMES[l]}: {q1.PLACEMENT_NAMES[r['placement']]} tpb={r['tpb']} tpt={r['tpt']} "
              f"{r['kernel_ms']:.3f} ms {r['gbps']:.1f} GB/s", flush=True)
    print(f"autotune: {len(results)} scan shapes, {len(batch_results)} batch shapes in {time.time() - t0:.1f}s",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
