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
    (["--cutoff=1997-02-29"], "Cannot parse the date"),
])
def test_bad_options_exit_with_failure(args, msg):
    r = run(args)
    assert r.returncode == 1
    assert msg in r.stderr


# lineitem.tbl lines in dbgen's text form (16 '|'-terminated fields; quantity printed without
# decimals, as dbgen does, or with them), including a leap day, the first and last ship
# dates of the generator's range and the value extremes.
DBGEN_LINES = [
    "1|155190|7706|1|17|21168.23|0.04|0.02|N|O|1996-03-13|1996-02-12|1996-03-22|DELIVER IN PERSON|TRUCK|egular courts above the|",
    "1|67310|7311|2|36|45983.16|0.09|0.06|N|O|1996-04-12|1996-02-28|1996-04-20|TAKE BACK RETURN|MAIL|ly final dependencies: slyly bold |",
    "3|4297|1798|1|45|54058.05|0.06|0.00|R|F|1994-02-02|1994-01-04|1994-02-23|NONE|AIR|ongside of the furiously brave acco|",
    "32|82704|7721|1|28|47227.60|0.05|0.08|N|O|1995-10-23|1995-08-27|1995-10-26|TAKE BACK RETURN|TRUCK|sleep quickly. req|",
    "35|1|1|1|24|22824.48|0.02|0.00|A|F|1992-02-29|1992-03-01|1992-03-02|NONE|AIR|a leap day|",
    "36|1|1|1|50.00|104999.50|0.10|0.08|R|F|1998-12-01|1998-12-01|1998-12-31|NONE|SHIP|the last ship date|",
    "37|1|1|1|1|900.00|0.00|0.00|A|F|1992-01-02|1992-01-02|1992-01-03|NONE|RAIL|the first ship date|",
    "38|1|1|1|2.5|1800.9|0.1|0.01|N|F|1995-06-17|1995-06-17|1995-06-18|NONE|FOB|one fraction digit|",
]


def expected_columns(lines):
    """The reference-layout columns (src/data_types.h:19-25) that the text fixes, computed
    from the characters of each line: decimals x100, MonetDB day numbers (1992-01-01 is
    727563, src/constants.hpp:166-186)."""
    import datetime
    from decimal import Decimal
    base = datetime.date(1992, 1, 1)
    cols = {k: [] for k in ("ship_date", "quantity", "extended_price", "discount", "tax", "return_flag",
                            "line_status")}
    for line in lines:
        f = line.split("|")
        for k, i in (("quantity", 4), ("extended_price", 5), ("discount", 6), ("tax", 7)):
            cols[k].append(int(Decimal(f[i]) * 100))
        cols["return_flag"].append(ord(f[8]))
        cols["line_status"].append(ord(f[9]))
        y, m, d = map(int, f[10].split("-"))
        cols["ship_date"].append(727563 + (datetime.date(y, m, d) - base).days)
    return cols


@pytest.mark.parametrize("eol", ["\n", "\r\n"])
def test_ingest_only_parses_dbgen_text(tmp_path, eol):
    """--ingest-only parses lineitem.tbl into the reference's column cache without a GPU;
    every column equals what the text of the lines fixes."""
    d = tmp_path / "data" / "1.000000"
    d.mkdir(parents=True)
    with open(d / "lineitem.tbl", "w", newline="") as f:
        f.write(eol.join(DBGEN_LINES) + eol)
    r = run(["--data-dir=%s" % (tmp_path / "data"), "--ingest-only"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "table length: %d records" % len(DBGEN_LINES) in r.stdout
    want = expected_columns(DBGEN_LINES)
    assert want["ship_date"][4] == 727563 + 59 and want["quantity"][7] == 250 and want["extended_price"][7] == 180090
    for name, col, dt in (("shipdate", "ship_date", np.int32), ("quantity", "quantity", np.int64),
                          ("extendedprice", "extended_price", np.int64), ("discount", "discount", np.int64),
                          ("tax", "tax", np.int64), ("returnflag", "return_flag", np.int8),
                          ("linestatus", "line_status", np.int8)):
        got = np.fromfile(d / f"{name}.bin", dt)
        assert got.tolist() == want[col], name


@pytest.mark.parametrize("bad,msg", [
    ("1|1|1|1|1x|900.00|0.00|0.00|A|F|1992-01-02|x|x|x|x|x|", "bad decimal"),
    ("1|1|1|1|1|900.001|0.00|0.00|A|F|1992-01-02|x|x|x|x|x|", "bad decimal"),
    ("1|1|1|1|1|900.00|.5|0.00|A|F|1992-01-02|x|x|x|x|x|", "bad decimal"),
    ("1|1|1|1|1|900.00|0.00|0.00|A|F|1995-02-29|x|x|x|x|x|", "bad ship date"),
    ("1|1|1|1|1|900.00|0.00|0.00|A|F|1992-1-02|x|x|x|x|x|", "bad ship date"),
    ("1|1|1|1|1|900.00|0.00|0.00|AF|F|1992-01-02|x|x|x|x|x|", "bad return flag"),
    ("1|1|1|1|1|900.00|0.00|0.00|A|F", "fewer than 11 fields"),
])
def test_ingest_only_rejects_malformed_lines(tmp_path, bad, msg):
    d = tmp_path / "data" / "1.000000"
    d.mkdir(parents=True)
    (d / "lineitem.tbl").write_text(DBGEN_LINES[0] + "\n" + bad + "\n")
    r = run(["--data-dir=%s" % (tmp_path / "data"), "--ingest-only"], cwd=tmp_path)
    assert r.returncode == 1
    assert msg in r.stderr and "line 2" in r.stderr
    assert not (d / "shipdate.bin").exists()


def test_ingest_only_needs_a_data_dir():
    r = run(["--ingest-only"])
    assert r.returncode == 1 and "--data-dir" in r.stderr


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
