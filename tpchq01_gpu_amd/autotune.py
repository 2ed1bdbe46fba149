"""Launch-shape autotuner for the Q1 kernels (replaces the reference's
scripts/try_grid_params.py and the hand-entered optimum table of
scripts/benchmark_with_optimal_params.py:39-55).

For every (layout, placement) it sweeps threads_per_block x tuples_per_thread, times the
kernel alone with HIP events (q1_last_timing().kernel_ms, the minimum over repetitions),
skips shapes that leave fewer than 2 blocks per compute unit (as try_grid_params.py:78-80
does), rotates between table copies when the table is smaller than the 256 MB Infinity
Cache so the timing reflects HBM, and writes

  tuning/gfx950.txt   "launch <layout> <placement> <tpb> <tpt>" and "best <layout> <placement>"
                      lines, read by q1_init (or Q1_TUNING_FILE)
  tuning/gfx950.json  every measurement

Usage: python -m tpchq01_gpu_amd.autotune [--rows N] [--layouts compressed,...]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import PACKAGE_DIR, q1

THREADS = [64, 128, 256]
TUPLES = [128, 256, 512, 1024, 2048, 4096, 8192, 16384]
MALL_BYTES = 256 << 20


def sweep(ctx: q1.Context, layout: int, rows: int, placements, reps: int, seed: int = 0x51C00010,
          log=print):
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
            for tpb in THREADS:
                for tpt in TUPLES:
                    blocks = -(-rows // (tpb * tpt))
                    if blocks < 2 * cus:
                        continue
                    launch = q1.make_launch(tpb, tpt)
                    best = float("inf")
                    # the memory-bound-by-atomics placements are 20-100x slower: one run each
                    for r in range(1 if p in (q1.LOCAL_MEM, q1.GLOBAL) else reps):
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


def best_table(results):
    best = {}
    for r in results:
        k = (r["layout"], r["placement"])
        if k not in best or r["kernel_ms"] < best[k]["kernel_ms"]:
            best[k] = r
    per_layout = {}
    for (l, p), r in best.items():
        if l not in per_layout or r["kernel_ms"] < per_layout[l]["kernel_ms"]:
            per_layout[l] = r
    return best, per_layout


def write_tables(results, out_dir: str, arch: str = "gfx950"):
    os.makedirs(out_dir, exist_ok=True)
    best, per_layout = best_table(results)
    with open(os.path.join(out_dir, f"{arch}.txt"), "w") as f:
        f.write("# written by tpchq01_gpu_amd.autotune: launch <layout> <placement> <tpb> <tpt> / best <layout> <placement>\n")
        for (l, p), r in sorted(best.items()):
            f.write(f"launch {l} {p} {r['tpb']} {r['tpt']}\n")
        for l, r in sorted(per_layout.items()):
            f.write(f"best {l} {r['placement']}\n")
    with open(os.path.join(out_dir, f"{arch}.json"), "w") as f:
        json.dump({"measurements": results,
                   "best": [{**r, "layout": q1.LAYOUT_NAMES[l], "placement": q1.PLACEMENT_NAMES[p]}
                            for (l, p), r in sorted(best.items())]}, f, indent=1)
    return best, per_layout


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=600_000_000)
    ap.add_argument("--layouts", default="compressed")
    ap.add_argument("--placements", default="in_registers,in_registers_per_thread,local_mem,shared_mem_per_thread,global")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(PACKAGE_DIR, "tuning"))
    args = ap.parse_args(argv)
    names = {"reference": q1.REFERENCE, "uncompressed": q1.UNCOMPRESSED, "compressed": q1.COMPRESSED,
             "compressed_pushdown": q1.COMPRESSED_PUSHDOWN}
    layouts = [names[x] for x in args.layouts.split(",")]
    placements = [q1.PLACEMENTS[x] for x in args.placements.split(",")]
    t0 = time.time()
    with q1.Context(0) as ctx:
        results = []
        for l in layouts:
            results += sweep(ctx, l, args.rows, placements, args.reps, log=lambda s: print(s, flush=True))
    best, per_layout = write_tables(results, args.out)
    for l, r in per_layout.items():
        print(f"best {q1.LAYOUT_NAMES[l]}: {q1.PLACEMENT_NAMES[r['placement']]} tpb={r['tpb']} tpt={r['tpt']} "
              f"{r['kernel_ms']:.3f} ms {r['gbps']:.1f} GB/s", flush=True)
    print(f"autotune: {len(results)} shapes in {time.time() - t0:.1f}s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
