"""The tpch_01 host driver: the reference's command-line contract (src/main.cu:255-360) on
the CPU, and end-to-end runs against the oracle on the GPU (synthetic rows, the .tbl
ingest and the .bin column cache)."""
import os
import re
import subprocess

import numpy as np
import pytest

from tpchq01_gpu_amd import BIN_DIR, q1

DRIVER = os.path.join(BIN_DIR, "tpch_01")


def run(args, cwd=None, timeout=300):
    return subprocess.run([DRIVER] + args, capture_output=True, text=True, cwd=cwd, timeout=timeout)


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "Usage"),
    (["notanoption"], "Usage"),
    (["--use-filter-pushdown"], "only currently supported when compression is applied"),
    (["--hash-table-placement=hashy"], 'No kernel variant named "hashy"'),
    (["--threads-per-block=100"], "64, 128 or 256"),
    (["--threads-per-block=192"], "64, 128 or 256"),
    (["--runs=0"], "Number of runs must be positive"),
    (["--cpu-fraction=1.5"], "range 0.0 - 1.0"),
    (["--sf=0"], "Invalid scale factor"),
    (["--streams=0"], "--streams"),
    (["--tuples-per-kernel-launch=100"], "multiple of 32"),
    (["--cutoff=1998-13-01"], "Cannot parse the date"),
])
def test_bad_options_exit_with_failure(args, msg):
    r = run(args)
    assert r.returncode == 1
    assert msg in r.stderr


def parse_printed(out: str):
    groups = {}
    for m in re.finditer(r"\| # (\w) \| (\w)\s+\|\s+(-?[\d.]+) \|\s+(-?[\d.]+) \|\s+(-?[\d.]+) \|\s+(-?[\d.]+) \|\s+(\d+) \|", out):
        rf, ls, qty, price, dp, ch, cnt = m.groups()
        num = lambda s: int(s.replace(".", ""))
        groups[rf + ls] = dict(sum_qty=num(qty), sum_base_price=num(price), sum_disc_price=num(dp),
                               sum_charge=num(ch), count=int(cnt))
    return groups


def expected(oracle, ref, rows, cutoff=729999):
    want, _ = oracle.naive(ref, 0, rows, cutoff)
    out = {}
    for g, v in enumerate(want):
        if v["count"]:
            out["ANR"[g >> 1] + "FO"[g & 1]] = {k: v[k] for k in ("sum_qty", "sum_base_price", "sum_disc_price",
                                                                 "sum_charge", "count")}
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [
    [],
    ["--apply-compression"],
    ["--apply-compression", "--use-filter-pushdown"],
    ["--apply-compression", "--use-coprocessing", "--cpu-fraction=0.3"],
    ["--use-coprocessing"],
    ["--apply-compression", "--device-resident", "--hash-table-placement=per-thread-shared-mem"],
    ["--device-resident", "--hash-table-placement=local-mem"],
    ["--apply-compression", "--use-filter-pushdown", "--device-resident", "--use-coprocessing"],
    ["--apply-compression", "--hash-table-placement=global", "--streams=2", "--tuples-per-kernel-launch=262144"],
])
def test_driver_end_to_end(oracle, tmp_path, flags):
    rows = 1_500_017
    r = run(["--rows=%d" % rows, "--runs=2", "--print-results", "--seed=0x51C00020"] + flags, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    ref = q1.generate_host(q1.REFERENCE, 0x51C00020, rows)
    got = parse_printed(r.stdout)
    assert got == expected(oracle, ref, rows)
    assert len(open(tmp_path / "results.csv").read().split()) == 2


def write_tbl(path, ref, rows):
    """lineitem.tbl in dbgen's format (16 '|'-terminated fields; 5-11 are the Q1 columns)."""
    import datetime
    base = datetime.date(1992, 1, 1).toordinal()
    with open(path, "w") as f:
        for i in range(rows):
            d = datetime.date.fromordinal(int(ref["ship_date"][i]) - 727563 + base)
            q, p = int(ref["quantity"][i]), int(ref["extended_price"][i])
            disc, tax = int(ref["discount"][i]), int(ref["tax"][i])
            f.write(f"{i}|1|1|1|{q // 100}.{q % 100:02d}|{p // 100}.{p % 100:02d}|0.{disc:02d}|0.{tax:02d}|"
                    f"{chr(ref['return_flag'][i])}|{chr(ref['line_status'][i])}|{d.isoformat()}|"
                    f"1998-01-01|1998-01-02|NONE|AIR|comment|\n")


@pytest.mark.gpu
def test_driver_tbl_ingest_and_bin_cache(oracle, tmp_path):
    rows = 20_011
    ref = q1.generate_host(q1.REFERENCE, 0x51C00021, rows)
    d = tmp_path / "data" / "1.000000"
    d.mkdir(parents=True)
    write_tbl(d / "lineitem.tbl", ref, rows)
    want = expected(oracle, ref, rows)
    r = run(["--data-dir=%s" % (tmp_path / "data"), "--runs=1", "--print-results", "--apply-compression"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "CSV read & parsed" in r.stdout
    assert parse_printed(r.stdout) == want
    # the second run reads the binary column cache written by the first
    assert (d / "shipdate.bin").stat().st_size == 4 * rows
    assert np.array_equal(np.fromfile(d / "quantity.bin", np.int64), ref["quantity"])
    r = run(["--data-dir=%s" % (tmp_path / "data"), "--runs=1", "--print-results"], cwd=tmp_path)
    assert r.returncode == 0 and "CSV read" not in r.stdout
    assert parse_printed(r.stdout) == want


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [["--apply-compression"], ["--apply-compression", "--device-resident"],
                                   ["--apply-compression", "--use-filter-pushdown", "--filter-once"], []])
def test_driver_tbl_outside_the_domain(oracle, tmp_path, flags):
    """A .tbl file with quantities of 99 units and prices of 2^24 cents or more (outside the
    TPC-H domain the narrow kernels assume): the driver checks the ingested columns, runs
    the 64-bit kernels, and prints the oracle's result."""
    rows = 20_011
    ref = q1.generate_host(q1.REFERENCE, 0x51C00022, rows)
    ref["quantity"][::7] = 9900
    ref["extended_price"][3::11] = 1 << 25
    d = tmp_path / "data" / "1.000000"
    d.mkdir(parents=True)
    write_tbl(d / "lineitem.tbl", ref, rows)
    r = run(["--data-dir=%s" % (tmp_path / "data"), "--runs=2", "--print-results"] + flags, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "Values outside the TPC-H domain" in r.stdout
    assert parse_printed(r.stdout) == expected(oracle, ref, rows)


@pytest.mark.gpu
def test_driver_filter_once(oracle, tmp_path):
    """--filter-once builds the pushdown bitmap before the timed runs (host and device)."""
    rows = 1_000_003
    ref = q1.generate_host(q1.REFERENCE, 0x51C00023, rows)
    for extra in ([], ["--device-resident"], ["--device-resident", "--use-coprocessing"]):
        r = run(["--rows=%d" % rows, "--runs=2", "--print-results", "--seed=0x51C00023", "--apply-compression",
                 "--use-filter-pushdown", "--filter-once"] + extra, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        assert "Filter bitmap built once" in r.stdout
        assert parse_printed(r.stdout) == expected(oracle, ref, rows), extra


@pytest.mark.gpu
def test_driver_device_info():
    r = run(["--device"])
    assert r.returncode == 1 and "gfx950" in r.stdout
