"""Host-side product code against the oracle, on the CPU (no GPU needed): the seeded
generator, the column compression (src/main.cu:773-803), the pushdown producer
(src/main.cu:362-389), the CPU coprocessor aggregation and the date numbering."""
import datetime

import numpy as np
import pytest

from tpchq01_gpu_amd import q1

SEED = 0x51C00000


def as_ref_layout(cols: dict) -> dict:
    """UNCOMPRESSED columns widened to the reference layout (same values)."""
    return {
        "ship_date": cols["ship_date"].astype(np.int32),
        "quantity": cols["quantity"].astype(np.int64),
        "extended_price": cols["extended_price"].astype(np.int64),
        "discount": cols["discount"].astype(np.int64),
        "tax": cols["tax"].astype(np.int64),
        "return_flag": cols["return_flag"],
        "line_status": cols["line_status"],
    }


def as_dicts(groups):
    return [dict(sum_qty=g.sum_qty, sum_base_price=g.sum_base_price, sum_disc_price=g.sum_disc_price,
                 sum_charge=g.sum_charge, sum_disc=g.sum_disc, count=g.count) for g in groups]


def test_day_number_matches_oracle(oracle):
    d = datetime.date(1900, 1, 1)
    while d.year < 2100:
        assert q1.day_number(d.year, d.month, d.day) == oracle.todate(d.day, d.month, d.year)
        d += datetime.timedelta(days=29)
    assert q1.day_number(1998, 9, 2) == 729999


@pytest.mark.parametrize("rows", [0, 1, 31, 32, 33, 1000, 100_003])
def test_generator_layouts_agree(oracle, rows):
    ref = q1.generate_host(q1.REFERENCE, SEED, rows)
    unc = q1.generate_host(q1.UNCOMPRESSED, SEED, rows)
    comp = q1.generate_host(q1.COMPRESSED, SEED, rows)
    assert np.array_equal(as_ref_layout(unc)["quantity"], ref["quantity"])
    for k in ("ship_date", "extended_price", "return_flag", "line_status"):
        assert np.array_equal(unc[k].astype(np.int64), ref[k].astype(np.int64)), k
    for k in ("discount", "tax"):
        assert np.array_equal(unc[k].astype(np.int64), ref[k]), k
    if rows:
        want = oracle.compress(ref, rows)
        for k in want:
            assert np.array_equal(comp[k], want[k]), k


def test_generator_row_ranges_regenerate():
    full = q1.generate_host(q1.COMPRESSED, SEED, 4096)
    part = q1.generate_host(q1.COMPRESSED, SEED, 1024, row_begin=2048)
    for k in ("ship_date", "quantity", "extended_price", "discount", "tax"):
        assert np.array_equal(full[k][2048:3072], part[k]), k
    assert np.array_equal(full["return_flag"][128:192], part["return_flag"])
    assert np.array_equal(full["line_status"][64:96], part["line_status"])


def test_generator_distribution():
    n = 400_000
    ref = q1.generate_host(q1.REFERENCE, SEED + 1, n)
    ship = ref["ship_date"]
    assert ship.min() >= 727564 and ship.max() <= 730089
    assert set(np.unique(ref["discount"])) <= set(range(11))
    assert set(np.unique(ref["tax"])) <= set(range(9))
    q = ref["quantity"] // 100
    assert q.min() >= 1 and q.max() <= 50
    part = ref["extended_price"] // q
    assert np.all(ref["extended_price"] % q == 0) and part.min() >= 90000 and part.max() <= 209999
    sel = (ship <= 729999).mean()
    assert abs(sel - 0.9644) < 0.003          # uniform dates with a 90-day tail cut
    rf = ref["return_flag"].view("S1")
    ls = ref["line_status"].view("S1")
    m = ship <= 729999
    share = lambda a, b: ((rf == a) & (ls == b) & m).mean()
    assert abs(share(b"A", b"F") - 0.2469) < 0.004
    assert abs(share(b"R", b"F") - 0.2469) < 0.004
    assert abs(share(b"N", b"F") - 0.0061) < 0.001
    assert abs(share(b"N", b"O") - 0.4644) < 0.004
    assert share(b"A", b"O") == 0 and share(b"R", b"O") == 0   # receipt > ship


@pytest.mark.parametrize("rows", [1, 17, 64, 1000, 65_537])
def test_compress_matches_oracle(oracle, rows):
    ref = q1.generate_host(q1.REFERENCE, SEED + 2, rows)
    got = q1.compress_columns(ref, rows, threads=3)
    want = oracle.compress(ref, rows)
    for k in want:
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("rows", [1, 31, 32, 33, 1000, 99_999])
@pytest.mark.parametrize("cutoff", [729999, 729906, 730088, 727000])
def test_cpu_filter_matches_oracle(oracle, rows, cutoff):
    comp = q1.generate_host(q1.COMPRESSED, SEED + 3, rows)
    got = q1.cpu_build_filter(comp["ship_date"], cutoff, threads=4)
    assert np.array_equal(got, oracle.precompute_filter(comp["ship_date"], rows, cutoff))


@pytest.mark.parametrize("layout", [q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED, q1.COMPRESSED_PUSHDOWN])
@pytest.mark.parametrize("rng", [(0, 0), (0, 1), (5, 1000), (0, 65_536), (17, 200_000)])
def test_cpu_coprocessor_matches_oracle(oracle, layout, rng):
    begin, count = rng
    n = begin + count
    ref = q1.generate_host(q1.REFERENCE, SEED + 4, n)
    want, unexpected = oracle.naive(ref, begin, n, 729906)
    assert unexpected == 0
    if layout == q1.REFERENCE:
        cols = ref
    elif layout == q1.UNCOMPRESSED:
        cols = q1.generate_host(q1.UNCOMPRESSED, SEED + 4, n)
    else:
        cols = q1.generate_host(q1.COMPRESSED, SEED + 4, n)
        if layout == q1.COMPRESSED_PUSHDOWN:
            cols = q1.pushdown_columns(cols, 729906, n)
    got = q1.cpu_run(layout, cols, begin, count, 729906, threads=4)
    assert as_dicts(got) == want


def test_cpu_coprocessor_out_of_domain(oracle):
    """Values far outside TPC-H (negative, huge): exact int64/int128 semantics."""
    import pinning
    rng = np.random.default_rng(1)
    rows = pinning.random_rows(rng, 3000, tpch=False)
    cols = pinning.ref_columns(rows)
    want, _ = oracle.naive(cols, 0, 3000, 729999)
    assert as_dicts(q1.cpu_run(q1.REFERENCE, cols, 0, 3000, 729999, threads=3)) == want


def test_result_merge_is_exact():
    a, b, out = q1.Result(), q1.Result(), q1.Result()
    a.g[0].sum_charge_lo = 2**64 - 5
    a.g[0].sum_charge_hi = 1
    b.g[0].sum_charge_lo = 10
    b.g[0].sum_charge_hi = -3
    a.g[1].sum_disc_price_lo = 1
    b.g[1].sum_disc_price_lo = 2**64 - 1
    b.g[1].sum_disc_price_hi = -1
    q1.lib.q1_result_merge(a, b, out)
    r = q1.result_to_list(out)
    assert r[0].sum_charge == (2**64 + 2**64 - 5) + (-3 * 2**64 + 10)
    assert r[1].sum_disc_price == 1 + (-(2**64) + 2**64 - 1)


def test_shipped_tuning_table_is_well_formed():
    """The autotuning table that q1_init loads next to the library: every launch shape is
    one the kernels accept, every layout names a best placement, and the batched kernel's
    shapes (compressed layout, both batch placements) and the streaming pipeline's chunk
    rows and stream counts are ones q1_init accepts."""
    import os

    from tpchq01_gpu_amd import PACKAGE_DIR
    path = os.path.join(PACKAGE_DIR, "tuning", "gfx950.txt")
    launches, best, batch, at = {}, {}, {}, {}
    for line in open(path):
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        if f[0] == "launch":
            l, p, tpb, tpt = map(int, f[1:5])
            assert tpb in (64, 128, 256) and tpt % 16 == 0 and 16 <= tpt <= 16384, line
            launches[(l, p)] = (tpb, tpt)
        elif f[0] == "launch_at":
            n, l, p, tpb, tpt = map(int, f[1:6])
            assert n > 0 and tpb in (64, 128, 256) and tpt % 16 == 0 and 16 <= tpt <= 16384, line
            at.setdefault(n, {})[(l, p)] = (tpb, tpt)
        elif f[0] == "batch":
            l, p, per_cu, lane_rows = map(int, f[1:5])
            most = 256 if p == q1.PER_THREAD_SHARED_MEM else 16384
            assert 1 <= per_cu <= 64 and lane_rows % 16 == 0 and 16 <= lane_rows <= most, line
            batch[(l, p)] = (per_cu, lane_rows)
        elif f[0] == "streaming":
            l, chunk, streams = map(int, f[1:4])
            assert 0 <= l <= 3 and chunk > 0 and chunk % 32 == 0 and 1 <= streams <= 16, line
        else:
            assert f[0] == "best", line
            best[int(f[1])] = int(f[2])
    layouts = (q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED, q1.COMPRESSED_PUSHDOWN)
    assert set(best) == set(layouts)
    for l in layouts:
        assert (l, best[l]) in launches
    assert set(batch) == {(q1.COMPRESSED, q1.IN_REGISTERS), (q1.COMPRESSED, q1.PER_THREAD_SHARED_MEM)}
    # shapes at several row counts: each count covers every tuned (layout, placement), and the
    # largest count's shapes are the "launch" lines
    if at:
        for n, shapes in at.items():
            assert set(shapes) == set(launches), n
        assert at[max(at)] == launches


def _compressed_cols(ship, qty, price, disc, tax, rf, ls):
    """Compressed host columns from per-row arrays (rf codes 0..3, ls bits)."""
    n = len(ship)
    cols = q1.empty_columns(q1.COMPRESSED, n)
    cols["ship_date"][:] = ship
    cols["quantity"][:] = qty
    cols["extended_price"][:] = price
    cols["discount"][:] = disc
    cols["tax"][:] = tax
    for i in range(n):
        cols["return_flag"][i // 16] |= np.uint32(int(rf[i]) << (2 * (i % 16)))
        cols["line_status"][i // 32] |= np.uint32(int(ls[i]) << (i % 32))
    return cols


def test_cpu_coprocessor_compressed_extremes_cross_int64():
    """Compressed columns at their type limits (u8 discount, tax, quantity 255, u32 price
    2^32-1) in one group over 40,000 rows: the charge sum passes -2^63, so the int64 runs of
    2^14 rows must widen exactly.  Every row is the same, so each sum is the row count times
    the row's value."""
    n = 40_000
    z = np.zeros(n, dtype=np.int64)
    cols = _compressed_cols(z, z + 255, z + (2**32 - 1), z + 255, z + 255, z, z)
    g = q1.cpu_run(q1.COMPRESSED, cols, 0, n, 727563, threads=3)[0]
    disc_price = (100 - 255) * (2**32 - 1)
    assert g.count == n
    assert g.sum_qty == n * 255 * 100
    assert g.sum_base_price == n * (2**32 - 1)
    assert g.sum_disc == n * 255
    assert g.sum_disc_price == n * disc_price
    assert g.sum_charge == n * disc_price * (255 + 100)
    assert g.sum_charge < -(2**63)


def test_cpu_coprocessor_compressed_full_range_random():
    """Random values over the compressed columns' whole ranges (rf codes 0..3, so some rows
    fall outside the six groups), 50,000 rows, against Python integer arithmetic."""
    rng = np.random.default_rng(5)
    n = 50_000
    ship = rng.integers(0, 2**16, n)
    qty, disc, tax = (rng.integers(0, 256, n) for _ in range(3))
    price = rng.integers(0, 2**32, n)
    rf, ls = rng.integers(0, 4, n), rng.integers(0, 2, n)
    cols = _compressed_cols(ship, qty, price, disc, tax, rf, ls)
    cut = 40_000
    got = q1.cpu_run(q1.COMPRESSED, cols, 0, n, 727563 + cut, threads=4)
    want = [dict(sum_qty=0, sum_base_price=0, sum_disc_price=0, sum_charge=0, sum_disc=0, count=0) for _ in range(6)]
    for i in range(n):
        slot = int(rf[i]) << 1 | int(ls[i])
        if int(ship[i]) > cut or slot >= 6:
            continue
        w = want[slot]
        dp = (100 - int(disc[i])) * int(price[i])
        w["sum_qty"] += int(qty[i]) * 100
        w["sum_base_price"] += int(price[i])
        w["sum_disc"] += int(disc[i])
        w["sum_disc_price"] += dp
        w["sum_charge"] += dp * (100 + int(tax[i]))
        w["count"] += 1
    assert as_dicts(got) == want


@pytest.mark.parametrize("layout", [q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED])
def test_host_domain_check(layout):
    """q1_host_columns_in_tpch_domain: generated rows lie inside the TPC-H value domain; one
    row beyond a bound of any of the four measured columns (discount or tax above 15,
    quantity above 63 units, price at 2^24 or more, a negative discount) puts the columns
    outside it."""
    import ctypes as C

    def check(cols, n):
        flag = C.c_int(-1)
        st = q1.lib.q1_host_columns_in_tpch_domain(layout, C.byref(q1.columns_struct(cols)), n, 3, C.byref(flag))
        assert st == q1.OK
        return flag.value

    n = 10_007
    base = q1.generate_host(layout, SEED, n)
    assert check(base, n) == 1
    assert check(base, 0) == 1
    qscale = 1 if layout == q1.COMPRESSED else 100
    for col, value in (("discount", 16), ("tax", 16), ("quantity", 64 * qscale), ("extended_price", 1 << 24)):
        for row in (0, n // 2, n - 1):
            cols = {k: v.copy() for k, v in base.items()}
            cols[col][row] = value
            assert check(cols, n) == 0, (col, row)
            # the row past the range is not looked at
            assert check(cols, row) == 1, (col, row)
    if layout != q1.COMPRESSED:
        cols = {k: v.copy() for k, v in base.items()}
        cols["discount"][7] = -1
        assert check(cols, n) == 0
