#!/usr/bin/env python3
"""TPC-H Q1 scan -> filter -> group-aggregate on one MI355X.

A step is one pass of Q1 over the whole device-resident table (600M rows, compressed
layout: 9.375 B/row), launched through the C ABI's q1_enqueue (the
drop-in kernel launch) into one device aggregate table.  Timed region: K back-to-back
passes between two synchronisations; value = rows * K / seconds.  The kernel's average
duration comes from HIP events recorded on the library's stream around the K launches.

Prints one JSON line (see the fields below).  --gpus N > 1 exits: this path runs on one
GPU here (a multi-GPU run would shard by row range and exchange a 384-byte table).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rows/s + % HBM roofline at 600M rows; batched req/s and p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=600_000_000)
    ap.add_argument("--layout", default="compressed", choices=["compressed", "uncompressed", "reference"])
    ap.add_argument("--placement", default="auto")
    ap.add_argument("--threads-per-block", type=int, default=0)
    ap.add_argument("--tuples-per-thread", type=int, default=0)
    ap.add_argument("--cpu-rows", type=int, default=60_000_000, help="CPU baseline sample (reference layout)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-passes", type=int, default=25,
                    help="CPU baseline runs per kernel, the first (cold) one discarded (cpu/q1.cpp:268-299)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "q1_compressed_600M_traffic.md"))
    ap.add_argument("--batch-requests", type=int, default=256, help="requests in the batch leg (0 skips it)")
    ap.add_argument("--no-stream-probe", dest="stream_probe", action="store_false")
    ap.add_argument("--no-layouts", dest="layouts", action="store_false",
                    help="skip the pushdown / uncompressed / reference layouts leg")
    ap.add_argument("--no-placements", dest="placements", action="store_false",
                    help="skip the leg that runs the headline table through every placement")
    ap.add_argument("--no-host", dest="host", action="store_false",
                    help="skip the host-resident streaming and CPU/GPU split leg")
    ap.add_argument("--serve-plan", default="2000:60,8000:5",
                    help="serving leg: offered rate (req/s) and seconds, comma-separated pairs ('' skips it)")
    ap.add_argument("--copy-probe-bytes", type=int, default=0,
                    help="after the timed region, copy this many bytes twice with a plain device copy "
                         "kernel (scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profile(path: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled)."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(q1, args, batch_spec):
    """The oracle's restatements of the reference CPU kernels, timed on this host and used
    as the checker of the device results (the only place the bench touches oracle/).

    Protocol of cpu/q1.cpp:268-299: `args.cpu_passes` runs (default 25), the first (cold)
    discarded, the mean of the rest.  Timed, on the reference layout (38 B/row) of the same
    seeded generator the device reads:
      * 6,000,000 rows (the reference's CPU case): KernelNaive (cpu/kernel_naive.hpp) and
        KernelHyPer<true> (cpu/kernel_hyper.hpp) on one thread and morsel-parallel on
        `args.cpu_threads` threads, KernelX100 (cpu/kernel_x100.hpp) on both;
      * 60,000,000 rows: KernelNaive morsel-parallel;
      * the whole 600,000,000-row scan workload: KernelNaive and KernelHyPer
        morsel-parallel (the reported value is KernelNaive's rate here).
    The library is the oracle built at the highest ISA level the host supports (the
    reference builds with -march=native).  Returns the baseline and the oracle's groups for
    the 600M scan, the 60M sample and every request of the batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_lib import Oracle, baseline_build

    from tpchq01_gpu_amd import workloads
    path, flags = baseline_build()
    o = Oracle(path)
    threads, passes = args.cpu_threads, max(2, args.cpu_passes)
    cut = 729999

    def timed(rows, fn):
        times = []
        for _ in range(passes):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        hot = times[1:]
        return rows / (sum(hot) / len(hot))

    out = {"unit": "rows/s", "cores": threads, "kind": "port", "compiler_flags": flags,
           "passes": passes, "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    checks = {}
    # 6M rows: the reference's CPU case, every kernel on one thread and on all threads
    n = 6_000_000
    cols = q1.generate_host(q1.REFERENCE, workloads.SAMPLE_SEED, n, 0, threads)
    comp = o.compact(cols, n)
    out["rows_6M"] = {
        "naive_1_thread": timed(n, lambda: o.naive(cols, 0, n, cut)),
        "hyper_1_thread": timed(n, lambda: o.hyper(cols, 0, n, cut)),
        "x100_1_thread": timed(n, lambda: o.x100(comp, 0, n, cut, 1)),
        "naive": timed(n, lambda: o.naive_parallel(cols, 0, n, cut, threads)),
        "hyper": timed(n, lambda: o.hyper_parallel(cols, 0, n, cut, threads)),
        "x100": timed(n, lambda: o.x100(comp, 0, n, cut, threads)),
    }
    del cols, comp
    log(f"bench: cpu 6M {out['rows_6M']}")
    # 60M rows: the every-placement sample
    n = args.cpu_rows
    cols = q1.generate_host(q1.REFERENCE, workloads.SAMPLE_SEED, n, 0, threads)
    out["rows_60M"] = {"naive": timed(n, lambda: o.naive_parallel(cols, 0, n, cut, threads))}
    checks["sample"] = o.naive_parallel(cols, 0, n, cut, threads)[0]
    del cols
    log(f"bench: cpu 60M {out['rows_60M']}")
    # the whole scan workload
    n = args.rows
    cols = q1.generate_host(q1.REFERENCE, workloads.SCAN_SEED, n, 0, threads)
    naive = timed(n, lambda: o.naive_parallel(cols, 0, n, cut, threads))
    hyper = timed(n, lambda: o.hyper_parallel(cols, 0, n, cut, threads))
    groups, flag = o.hyper_parallel(cols, 0, n, cut, threads)
    checks["scan"] = o.naive_parallel(cols, 0, n, cut, threads)[0]
    checks["scan_hyper_agrees"] = flag == 0 and groups == checks["scan"]
    del cols
    out.update({"value": naive, "ms_per_pass": n / naive * 1e3,
                "sample": f"the whole scan workload: {n} rows in the reference layout (38 B/row), oracle "
                          f"KernelNaive restatement morsel-parallel on {threads} threads, mean of "
                          f"{passes - 1} passes after a cold one",
                "rows_600M": {"naive": naive, "hyper": hyper}})
    log(f"bench: cpu 600M {out['rows_600M']}")
    # the batch's requests, each on its own rows (checker only)
    if batch_spec:
        checks["batch"] = []
        for seed, rows, cutoff in batch_spec:
            cols = q1.generate_host(q1.REFERENCE, seed, rows, 0, threads)
            checks["batch"].append(o.naive_parallel(cols, 0, rows, cutoff, threads)[0])
    return out, checks


def hbm_probe(rows: int):
    """Practical HBM read ceiling: tpchq01_gpu_amd/bin/hbm_probe reads the same nine
    compressed columns (and the same bytes as one flat buffer) with no arithmetic, in a
    child process started before this one touches the GPU."""
    import subprocess
    exe = os.path.join(ROOT, "tpchq01_gpu_amd", "bin", "hbm_probe")
    r = subprocess.run([exe, str(rows)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        log(f"bench: hbm_probe failed ({r.returncode}): {r.stderr.strip()}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def batch_leg(q1, ctx, spec, headline=None):
    """Batch leg: len(spec) requests, each on its own seeded column slice (log-uniform
    1e4..2e7 rows, cutoff uniform in 1998-06-01..1998-11-30), in one q1_run_batch
    (persistent kernel replayed from a hipGraph) vs the plain per-request q1_run one at a
    time.  With `headline` (table, rows) it also runs that whole table as a batch of one
    request beside q1_run, which isolates the batched kernel's own cost.  Returns the leg's
    figures and the batch's results (checked against the oracle later, in the CPU baseline
    leg)."""
    import numpy as np
    n = len(spec)
    sizes = np.array([s[1] for s in spec], dtype=np.int64)
    cutoffs = np.array([s[2] for s in spec], dtype=np.int64)
    tables = [ctx.generate(q1.COMPRESSED, spec[r][0], int(sizes[r])) for r in range(n)]
    reqs = (q1.Request * n)(*[q1.Request(tables[r].handle, 0, int(sizes[r]), int(cutoffs[r]), 0) for r in range(n)])
    outs = (q1.Result * n)()
    # Build the graph, then warm: over the first ten or so back-to-back batches the device
    # time rises and falls again by 4-6 %, whatever ran before.
    for _ in range(8):
        ctx.check(q1.lib.q1_run_batch(ctx.handle, reqs, n, None, outs), "q1_run_batch")
    reps = 20
    dev = []
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.check(q1.lib.q1_run_batch(ctx.handle, reqs, n, None, outs), "q1_run_batch")
        dev.append(ctx.last_timing().kernel_ms)
    batched = (time.perf_counter() - t0) / reps
    batched_dev = sum(dev) / len(dev) * 1e-3
    tm = ctx.last_timing()
    shape = {"blocks": tm.num_blocks, "tuples_per_thread": tm.tuples_per_thread,
             "placement": q1.PLACEMENT_NAMES[tm.placement]}
    one_table = None
    if headline is not None:
        table, trows = headline
        req1 = (q1.Request * 1)(q1.Request(table.handle, 0, trows, 729999, 0))
        out1 = (q1.Result * 1)()
        bms, sms = [], []
        for rep in range(6):
            ctx.check(q1.lib.q1_run_batch(ctx.handle, req1, 1, None, out1), "q1_run_batch")
            b = ctx.last_timing().kernel_ms
            ctx.run_raw(table, 0, trows, 729999, q1.AUTO)
            if rep:
                bms.append(b)
                sms.append(ctx.last_timing().kernel_ms)
        one_table = {"rows": trows, "batch_kernel_ms": sum(bms) / len(bms), "scan_kernel_ms": sum(sms) / len(sms)}
        # the 256-request graph is rebuilt on the next call of that size
    singles = []
    for r in range(n):   # warm the per-request path
        ctx.run_raw(tables[r], 0, int(sizes[r]), int(cutoffs[r]), q1.AUTO)
    t0 = time.perf_counter()
    for r in range(n):
        singles.append(ctx.run_raw(tables[r], 0, int(sizes[r]), int(cutoffs[r]), q1.AUTO))
    one = time.perf_counter() - t0
    exact = all(q1.result_to_list(outs[r]) == q1.result_to_list(singles[r]) for r in range(n))
    results = [as_dicts(q1.result_to_list(outs[r])) for r in range(n)]
    for t in tables:
        t.free()
    rows = int(sizes.sum())
    return {"requests": n, "rows": rows, "batched_ms": batched * 1e3, "batched_device_ms": batched_dev * 1e3,
            "batched_device_ms_range": [min(dev), max(dev)],
            "batched_req_s": n / batched, "batched_rows_s": rows / batched,
            "batched_device_gbps": rows * q1.BITS_PER_ROW[q1.COMPRESSED] / 8 / batched_dev / 1e9,
            "one_at_a_time_ms": one * 1e3, "one_at_a_time_req_s": n / one,
            "speedup": one / batched, "exact_vs_single": exact, "shape": shape, "one_table": one_table}, results


def probe_ceiling(q1, probe, layout):
    """The probe's best pure-read rate for a layout's columns (GB/s), or None: it reads the
    compressed, uncompressed and reference columns with their kernels' access patterns."""
    if not probe:
        return None
    key = {q1.COMPRESSED: "best_columns_gbps", q1.UNCOMPRESSED: "best_uncompressed_gbps",
           q1.REFERENCE: "best_reference_gbps"}.get(layout)
    return probe.get(key) if key else None


def layouts_leg(q1, ctx, comp_table, rows: int, seed: int, reps: int = 5, probe=None):
    """The same rows in the other layouts (pushdown bitmap built on the device from the
    compressed table; uncompressed and reference layouts generated from the same seed):
    kernel time of q1_run with the tuned placement, mean of `reps` runs after one warm-up,
    and a check that every layout returns the compressed table's result exactly."""
    want = q1.result_to_list(ctx.run_raw(comp_table, 0, rows, 729999, q1.AUTO))
    out, exact = {}, True
    for name, make in (("compressed_pushdown", lambda: ctx.build_filter(comp_table, 729999)),
                       ("uncompressed", lambda: ctx.generate(q1.UNCOMPRESSED, seed, rows)),
                       ("reference", lambda: ctx.generate(q1.REFERENCE, seed, rows))):
        t = make()
        got = q1.result_to_list(ctx.run_raw(t, 0, rows, 729999, q1.AUTO))
        exact = exact and got == want
        ms = []
        for _ in range(reps):
            ctx.run_raw(t, 0, rows, 729999, q1.AUTO)
            ms.append(ctx.last_timing().kernel_ms)
        tm = ctx.last_timing()
        ms = sum(ms) / len(ms)
        algo = rows * q1.BITS_PER_ROW[t.layout] // 8
        out[name] = {"bytes_per_row": q1.BITS_PER_ROW[t.layout] / 8, "kernel_ms": ms,
                     "gbps": algo / (ms * 1e-3) / 1e9, "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "rows_s": rows / (ms * 1e-3), "placement": q1.PLACEMENT_NAMES[tm.placement],
                     "threads_per_block": tm.threads_per_block, "tuples_per_thread": tm.tuples_per_thread}
        ceiling = probe_ceiling(q1, probe, t.layout)
        if ceiling:
            out[name]["read_ceiling_gbps"] = ceiling
            out[name]["frac_of_read_ceiling"] = out[name]["gbps"] / ceiling
        t.free()
    out["exact_vs_compressed"] = exact
    # the pushdown producer (q1_build_filter on the device: 2 B/row read, 1 bit/row written)
    # against what each query saves by reading the bitmap instead of the ship dates
    builds, comp_ms = [], []
    for _ in range(reps):
        pd = ctx.build_filter(comp_table, 729999)
        builds.append(ctx.last_timing().kernel_ms)
        pd.free()
        ctx.run_raw(comp_table, 0, rows, 729999, q1.AUTO)
        comp_ms.append(ctx.last_timing().kernel_ms)
    build_ms, comp_ms = sum(builds) / len(builds), sum(comp_ms) / len(comp_ms)
    saved = comp_ms - out["compressed_pushdown"]["kernel_ms"]
    out["filter_producer"] = {"kernel_ms": build_ms, "gbps": rows * (2 + 1 / 8) / (build_ms * 1e-3) / 1e9,
                              "compressed_kernel_ms": comp_ms, "saved_per_query_ms": saved,
                              "queries_to_repay": build_ms / saved if saved > 0 else None}
    return out


def as_dicts(groups):
    return [dict(sum_qty=g.sum_qty, sum_base_price=g.sum_base_price, sum_disc_price=g.sum_disc_price,
                 sum_charge=g.sum_charge, sum_disc=g.sum_disc, count=g.count) for g in groups]


def placements_leg(q1, ctx, table, rows: int, want, reps: int = 10):
    """The headline table (compressed, device-resident) through every placement with its
    tuned launch: kernel time of q1_run (mean of `reps` after one warm-up), HBM fraction,
    and each placement's result against the headline result."""
    algo = rows * q1.BITS_PER_ROW[table.layout] // 8
    out = {}
    for p in (q1.IN_REGISTERS, q1.IN_REGISTERS_PER_THREAD, q1.LOCAL_MEM, q1.PER_THREAD_SHARED_MEM, q1.GLOBAL):
        got = q1.result_to_list(ctx.run_raw(table, 0, rows, 729999, p))
        ms = []
        for _ in range(reps):
            ctx.run_raw(table, 0, rows, 729999, p)
            ms.append(ctx.last_timing().kernel_ms)
        tm = ctx.last_timing()
        ms = sum(ms) / len(ms)
        out[q1.PLACEMENT_NAMES[p]] = {"kernel_ms": ms, "gbps": algo / (ms * 1e-3) / 1e9,
                                      "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "rows_s": rows / (ms * 1e-3),
                                      "threads_per_block": tm.threads_per_block,
                                      "tuples_per_thread": tm.tuples_per_thread, "exact_vs_headline": got == want}
    return out


def uncompressed_leg(q1, ctx, rows: int, seed: int, reps: int = 20):
    """`rows` device-resident uncompressed rows (20 B/row: int32 ship date and quantity,
    int64 price, int8 discount, tax, flags) through every placement: kernel time of q1_run
    with the placement's tuned launch, mean of `reps` runs after one warm-up.  The same
    seed and rows as the CPU baseline's sample, whose oracle result checks every placement."""
    t = ctx.generate(q1.UNCOMPRESSED, seed, rows)
    algo = rows * q1.BITS_PER_ROW[q1.UNCOMPRESSED] // 8
    out, results = {"rows": rows, "bytes_per_row": 20.0, "placements": {}}, []
    try:
        for p in (q1.IN_REGISTERS, q1.IN_REGISTERS_PER_THREAD, q1.LOCAL_MEM, q1.PER_THREAD_SHARED_MEM, q1.GLOBAL):
            name = q1.PLACEMENT_NAMES[p]
            got = ctx.run_raw(t, 0, rows, 729999, p)
            results.append(q1.result_to_list(got))
            ms = []
            for _ in range(reps):
                ctx.run_raw(t, 0, rows, 729999, p)
                ms.append(ctx.last_timing().kernel_ms)
            tm = ctx.last_timing()
            ms = sum(ms) / len(ms)
            out["placements"][name] = {"kernel_ms": ms, "rows_s": rows / (ms * 1e-3),
                                       "gbps": algo / (ms * 1e-3) / 1e9,
                                       "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                       "threads_per_block": tm.threads_per_block,
                                       "tuples_per_thread": tm.tuples_per_thread}
    finally:
        t.free()
    out["placements_agree"] = all(r == results[0] for r in results)
    out["_results"] = results
    return out


def host_leg(q1, ctx, table, rows: int, seed: int, want, cpu_threads: int, reps: int = 3):
    """Host-resident input, the reference's own mode: the same compressed rows generated
    into pinned host memory.  (1) q1_run_streaming: chunks of the autotuning table's rows
    per chunk copied over its number of streams, with the copies overlapping the kernels
    (src/main.cu:567-724), end-to-end time per query.  (2) q1_run_split, the CPU/GPU coprocessing split (src/main.cu:513-550):
    the device scans [0, split) of the resident table while `cpu_threads` host threads
    aggregate [split, rows) of the host copy, at the reference's default CPU fraction 0.5
    and at the fraction that balances the two sides' measured rates.  Every result is
    checked against the device-resident one."""
    cols = ctx.pinned_columns(q1.COMPRESSED, rows)
    try:
        t0 = time.perf_counter()
        q1.fill_host(q1.COMPRESSED, seed, cols, rows, 0, cpu_threads)
        gen_s = time.perf_counter() - t0
        nbytes = rows * q1.BITS_PER_ROW[q1.COMPRESSED] // 8
        chunk, streams = ctx.tuning_get_streaming(q1.COMPRESSED)
        dom = q1.make_launch(flags=q1.LAUNCH_ASSUME_TPCH_DOMAIN)   # generated rows lie in the TPC-H domain
        exact = ctx.run_streaming(q1.COMPRESSED, cols, rows, 729999, q1.AUTO, dom, chunk, streams) == want
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            got = ctx.run_streaming(q1.COMPRESSED, cols, rows, 729999, q1.AUTO, dom, chunk, streams)
            walls.append(time.perf_counter() - t0)
            exact = exact and got == want
        wall = sum(walls) / len(walls)
        out = {"rows": rows, "host_generate_s": gen_s,
               "streaming": {"chunk_rows": chunk, "streams": streams, "ms_per_query": wall * 1e3,
                             "rows_s": rows / wall, "h2d_gbps": nbytes / wall / 1e9, "exact": exact}}
        align = 1 << 21
        split_out = []
        gpu_rate = cpu_rate = None
        for frac in (0.5, None):
            if frac is None:   # balance the sides with the rates the 0.5 split measured
                frac = cpu_rate / (cpu_rate + gpu_rate)
            split = rows - int(rows * frac)
            split -= split % align
            got, cpu_ms, gpu_ms = ctx.run_split(table, cols, rows, frac, 729999, q1.AUTO, None, align, cpu_threads)
            ok = got == want
            t0 = time.perf_counter()
            for _ in range(reps):
                got, cpu_ms, gpu_ms = ctx.run_split(table, cols, rows, frac, 729999, q1.AUTO, None, align,
                                                    cpu_threads)
                ok = ok and got == want
            wall = (time.perf_counter() - t0) / reps
            gpu_rate = split / (gpu_ms * 1e-3) if gpu_ms > 0 else gpu_rate
            cpu_rate = (rows - split) / (cpu_ms * 1e-3) if cpu_ms > 0 and rows > split else cpu_rate
            split_out.append({"cpu_fraction": frac, "gpu_rows": split, "cpu_rows": rows - split,
                              "gpu_ms": gpu_ms, "cpu_ms": cpu_ms, "wall_ms": wall * 1e3,
                              "gpu_rows_s": split / (gpu_ms * 1e-3) if gpu_ms > 0 else None,
                              "cpu_rows_s": (rows - split) / (cpu_ms * 1e-3) if cpu_ms > 0 else None,
                              "cpu_threads": cpu_threads, "exact": ok})
        out["split"] = split_out
        return out
    finally:
        ctx.free_pinned_columns(cols)


def serving_leg(q1, ctx, plan):
    """Serving leg: Poisson arrivals over 8 HBM-resident compressed column sets
    (1M..600M rows), random row range and cutoff per request, through the
    continuous-batching server vs one-at-a-time dispatch of the plain kernel, for each
    (offered rate, seconds) of `plan`."""
    import numpy as np

    from tpchq01_gpu_amd import workloads
    from tpchq01_gpu_amd.serving import Server, make_request_stream, run_open_loop
    tables = [ctx.generate(q1.COMPRESSED, s, n) for s, n in workloads.serving_spec()]
    out = []
    for rate, seconds in plan:
        stream = make_request_stream(tables, rate, seconds, seed=9)
        # admission budget: three quarters of device memory (the server's default)
        with Server(ctx, max_batch=256, max_inflight=2, batch_window_us=0) as sv:
            done, wall = run_open_loop(sv, stream)
        lat = np.array([d[1] for d in done]) / 1e3
        bs = np.array([d[2] for d in done])
        ok = all(d[0] == q1.OK for d in done)
        # one at a time: the same arrivals served by q1_run in order on one thread
        t0 = time.perf_counter()
        one_lat = []
        for at, tab, b, nrows, c in stream:
            now = time.perf_counter() - t0
            if at > now:
                time.sleep(at - now)
            ctx.run_raw(tab, b, nrows, c, q1.AUTO)
            one_lat.append((time.perf_counter() - t0 - at) * 1e3)
        one_wall = time.perf_counter() - t0
        one_lat = np.array(one_lat)
        out.append({"rate": rate, "seconds": seconds, "requests": len(stream), "ok": ok,
                    "mean_rows": float(np.mean([r[3] for r in stream])),
                    "req_s": len(stream) / wall, "p50_ms": float(np.percentile(lat, 50)),
                    "p99_ms": float(np.percentile(lat, 99)), "mean_batch": float(bs.mean()),
                    "one_at_a_time": {"req_s": len(stream) / one_wall, "p50_ms": float(np.percentile(one_lat, 50)),
                                      "p99_ms": float(np.percentile(one_lat, 99))}})
        log(f"bench: serving {out[-1]}")
    for t in tables:
        t.free()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    probe = hbm_probe(args.rows) if args.stream_probe else None
    if probe:
        log(f"bench: read ceiling {probe['best_columns_gbps']:.0f} GB/s (columns), "
            f"{probe['best_flat_gbps']:.0f} GB/s (flat)")
    import torch

    from tpchq01_gpu_amd import q1

    layout = {"compressed": q1.COMPRESSED, "uncompressed": q1.UNCOMPRESSED, "reference": q1.REFERENCE}[args.layout]
    placement = q1.PLACEMENTS[args.placement]
    rows = args.rows
    ctx = q1.Context(0)
    info = ctx.device_info()
    log(f"bench: device {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")
    log(f"bench: generating {rows} rows ({q1.LAYOUT_NAMES[layout]}) on the device")
    table = ctx.generate(layout, 0x51C00002, rows)
    if placement == q1.AUTO:
        # the autotuning table's best placement for this layout
        ctx.run(table, 0, min(rows, 1 << 20), 729999, q1.AUTO)
        placement = ctx.last_timing().placement
    tuned = ctx.tuning_resolve(layout, placement, rows)
    launch = q1.make_launch(args.threads_per_block or tuned.threads_per_block,
                            args.tuples_per_thread or tuned.tuples_per_thread, 0,
                            q1.LAUNCH_ASSUME_TPCH_DOMAIN)
    log(f"bench: placement {q1.PLACEMENT_NAMES[placement]} tpb={launch.threads_per_block} "
        f"tpt={launch.tuples_per_thread}")

    # one reference pass through q1_run (fully reduced, host result)
    single = ctx.run_raw(table, 0, rows, 729999, placement, launch)
    single_ms = ctx.last_timing().kernel_ms
    log(f"bench: single q1_run kernel {single_ms:.3f} ms")

    dev_cols = table.device_columns()
    agg = torch.zeros(48, dtype=torch.int64, device="cuda")
    stream = torch.cuda.ExternalStream(ctx.stream())
    p_agg = C.c_void_p(agg.data_ptr())

    def enqueue():
        ctx.check(q1.lib.q1_enqueue(ctx.handle, placement, layout, C.byref(launch), p_agg, C.byref(dev_cols),
                                    0, rows, 729999, None), "q1_enqueue")

    for _ in range(args.warmup):
        enqueue()
    ctx.synchronize()
    with torch.cuda.stream(stream):
        agg.zero_()
    torch.cuda.synchronize()
    ctx.synchronize()

    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        enqueue()
    ev1.record(stream)
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps

    # K passes accumulate K times the single-pass table: a size-independent exactness check
    host = q1.DeviceAggregates()
    agg_host = agg.cpu().numpy()   # keep the host copy alive while it is read
    C.memmove(C.byref(host), agg_host.ctypes.data, C.sizeof(host))
    res = q1.Result()
    q1.lib.q1_aggregates_to_result(C.byref(host), C.byref(res))
    got, one = q1.result_to_list(res), q1.result_to_list(single)
    k = args.steps
    ok = all(g.count == k * s.count and g.sum_charge == k * s.sum_charge and
             g.sum_disc_price == k * s.sum_disc_price and g.sum_qty == k * s.sum_qty for g, s in zip(got, one))
    if not ok:
        log("bench: ERROR accumulated passes differ from K x single pass")
        sys.exit(1)

    if args.copy_probe_bytes:
        src = torch.ones(args.copy_probe_bytes // 8, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copy probe {args.copy_probe_bytes} bytes x2 done")
        del src, dst

    algo_bytes = rows * q1.BITS_PER_ROW[layout] // 8
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9
    traffic = traffic_from_profile(args.traffic_file) if layout == q1.COMPRESSED else None
    ceiling = probe_ceiling(q1, probe, layout)
    out = {
        "metric": METRIC,
        "value": rows * args.steps / elapsed,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded counter-based generator, dbgen value distributions)",
        "config": {"workload": f"TPC-H Q1, {rows} device-resident rows, {q1.LAYOUT_NAMES[layout]} layout",
                   "rows": rows, "layout": q1.LAYOUT_NAMES[layout],
                   "placement": q1.PLACEMENT_NAMES[placement],
                   "threads_per_block": launch.threads_per_block,
                   "tuples_per_thread": launch.tuples_per_thread,
                   "bytes_per_row": q1.BITS_PER_ROW[layout] / 8},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_ms": kernel_ms, "algorithmic_bytes": algo_bytes,
                     "read_ceiling_gbps": ceiling,
                     "frac_of_read_ceiling": achieved / ceiling if ceiling else None,
                     "copy_peak_gbps": probe.get("copy_gbps") if probe else None},
    }
    # The batch leg runs before the legs that allocate and free tens of GB (layouts,
    # placements, host-resident): after the host-resident leg its one-at-a-time loop of
    # q1_run calls measured up to 4x slower.
    spec, batch_results = [], None
    if args.batch_requests:
        from tpchq01_gpu_amd import workloads
        spec = workloads.batch_spec(args.batch_requests)
        log(f"bench: batch leg, {args.batch_requests} requests")
        out["batch"], batch_results = batch_leg(q1, ctx, spec,
                                                (table, rows) if layout == q1.COMPRESSED else None)
        # the same columns read by the batched kernel's resident-grid structure with no
        # arithmetic (hbm_probe's resident pattern): the batch's own ceiling
        res = probe.get("resident_gbps") if probe else None
        out["batch"]["resident_read_ceiling_gbps"] = res
        out["batch"]["frac_of_resident_read_ceiling"] = out["batch"]["batched_device_gbps"] / res if res else None
        log(f"bench: batch {out['batch']}")
    if args.layouts and layout == q1.COMPRESSED:
        log("bench: layouts leg")
        out["layouts"] = layouts_leg(q1, ctx, table, rows, 0x51C00002, probe=probe)
        # the same columns through the any-value kernel that runs outside the TPC-H value domain
        # (q1_enqueue without Q1_LAUNCH_ASSUME_TPCH_DOMAIN), same accumulate-and-compare check
        wide = q1.make_launch(256, 1024, 0, 0)
        with torch.cuda.stream(stream):
            agg.zero_()
        torch.cuda.synchronize()
        k = 5
        for _ in range(2):
            ctx.check(q1.lib.q1_enqueue(ctx.handle, q1.IN_REGISTERS_PER_THREAD, layout, C.byref(wide), p_agg,
                                        C.byref(dev_cols), 0, rows, 729999, None), "q1_enqueue")
        ctx.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            ctx.check(q1.lib.q1_enqueue(ctx.handle, q1.IN_REGISTERS_PER_THREAD, layout, C.byref(wide), p_agg,
                                        C.byref(dev_cols), 0, rows, 729999, None), "q1_enqueue")
        e1.record(stream)
        ctx.synchronize()
        torch.cuda.synchronize()
        wms = e0.elapsed_time(e1) / k
        host = q1.DeviceAggregates()
        agg_host = agg.cpu().numpy()
        C.memmove(C.byref(host), agg_host.ctypes.data, C.sizeof(host))
        res = q1.Result()
        q1.lib.q1_aggregates_to_result(C.byref(host), C.byref(res))
        exact = all(g.count == (k + 2) * s1.count and g.sum_charge == (k + 2) * s1.sum_charge
                    for g, s1 in zip(q1.result_to_list(res), one))
        out["layouts"]["compressed_without_domain_flag"] = {
            "kernel_ms": wms, "gbps": algo_bytes / (wms * 1e-3) / 1e9,
            "frac": algo_bytes / (wms * 1e-3) / 1e9 / HBM_PEAK_GBS, "threads_per_block": 256,
            "tuples_per_thread": 1024, "exact_vs_fast": exact}
        log(f"bench: layouts {out['layouts']}")
    if args.placements:
        log("bench: the headline table through every placement")
        out["placements"] = placements_leg(q1, ctx, table, rows, one)
        log(f"bench: placements {out['placements']}")
    if args.layouts and layout == q1.COMPRESSED:
        log(f"bench: {args.cpu_rows} uncompressed rows through every placement")
        out["uncompressed_60M"] = uncompressed_leg(q1, ctx, args.cpu_rows, 0x51C00000)
        log(f"bench: uncompressed_60M {out['uncompressed_60M']['placements']}")
    if args.host and layout == q1.COMPRESSED:
        log("bench: host-resident leg (streaming, CPU/GPU split)")
        out["host_resident"] = host_leg(q1, ctx, table, rows, 0x51C00002, one, args.cpu_threads)
        log(f"bench: host_resident {out['host_resident']}")
    plan = [(float(r), float(s)) for r, s in (x.split(":") for x in args.serve_plan.split(",") if x)]
    if plan:
        log(f"bench: serving leg, {plan} (req/s : seconds)")
        out["serving"] = serving_leg(q1, ctx, plan)
    table.free()
    ctx.close()
    out["cpu_baseline"] = None
    if not args.no_cpu:
        log(f"bench: CPU baseline ({args.cpu_passes} passes, {args.cpu_threads} threads) and oracle checks")
        out["cpu_baseline"], checks = cpu_baseline(q1, args, spec)
        # every device result above against the oracle's groups for the same rows
        out["exact_vs_oracle"] = as_dicts(one) == checks["scan"] and checks["scan_hyper_agrees"]
        if "uncompressed_60M" in out and out["uncompressed_60M"]["rows"] == args.cpu_rows:
            out["uncompressed_60M"]["exact_vs_oracle"] = all(
                as_dicts(r) == checks["sample"] for r in out["uncompressed_60M"]["_results"])
        if batch_results is not None:
            out["batch"]["exact_vs_oracle"] = batch_results == checks["batch"]
        log(f"bench: exact vs oracle: scan {out['exact_vs_oracle']}, "
            f"batch {out.get('batch', {}).get('exact_vs_oracle')}, "
            f"60M {out.get('uncompressed_60M', {}).get('exact_vs_oracle')}")
    if "uncompressed_60M" in out:
        out["uncompressed_60M"].pop("_results", None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
