"""Checks that pin the CPU oracle (oracle/q1_oracle.c) to what the mathematics and the
reference fix, independently of the oracle's own code:

* a calendar library routine (Python's proleptic Gregorian ordinal) and the reference's
  published day-number constants (src/constants.hpp:166-167);
* brute force on tiny tables against an independent formulation of the arithmetic (the
  expanded polynomials disc_price = 100p - dp and charge = 10000p + 100tp - 100dp - dtp
  in unbounded Python integers, grouping by the (returnflag, linestatus) characters);
* closed forms (n identical rows), the filter boundary, 128-bit sums crossing 2^63 and
  2^64 in both signs, invariants (counts and masked totals from numpy), the compression
  round trip with an independent bit packer, and the cutoff-shift invariance.

Each check returns True when the oracle passes.  tests/test_oracle.py requires all to
pass for the oracle and, for every planted-mistake build (-DQ1O_MUTANT=k), at least one
to fail.
"""
from __future__ import annotations

import datetime

import numpy as np

CUTOFF = 729999            # 1998-09-02, src/constants.hpp:167
FOR = 727563               # 1992-01-01, src/constants.hpp:166
RF, LS = "ANR", "FO"


def slot(rf: str, ls: str) -> int:
    return RF.index(rf) << 1 | LS.index(ls)


def py_q1(rows, cutoff):
    """Independent Q1 over python tuples (ship, qty, price, disc, tax, rf, ls)."""
    out = [dict(sum_qty=0, sum_base_price=0, sum_disc_price=0, sum_charge=0, sum_disc=0, count=0)
           for _ in range(6)]
    for ship, qty, price, disc, tax, rf, ls in rows:
        if ship <= cutoff:
            g = out[slot(rf, ls)]
            g["sum_qty"] += qty
            g["sum_base_price"] += price
            g["sum_disc_price"] += 100 * price - disc * price
            g["sum_charge"] += 10000 * price + 100 * tax * price - 100 * disc * price - disc * tax * price
            g["sum_disc"] += disc
            g["count"] += 1
    return out


def ref_columns(rows) -> dict:
    return {
        "ship_date": np.array([r[0] for r in rows], np.int32),
        "quantity": np.array([r[1] for r in rows], np.int64),
        "extended_price": np.array([r[2] for r in rows], np.int64),
        "discount": np.array([r[3] for r in rows], np.int64),
        "tax": np.array([r[4] for r in rows], np.int64),
        "return_flag": np.frombuffer("".join(r[5] for r in rows).encode(), np.uint8).copy(),
        "line_status": np.frombuffer("".join(r[6] for r in rows).encode(), np.uint8).copy(),
    }


def random_rows(rng, n, *, tpch=True, ship_lo=FOR + 1, ship_hi=730089):
    rows = []
    for _ in range(n):
        if tpch:
            qty = int(rng.integers(1, 51)) * 100
            price = qty // 100 * int(rng.integers(90000, 210000))
            disc, tax = int(rng.integers(0, 11)), int(rng.integers(0, 9))
        else:
            qty = int(rng.integers(-10**6, 10**6))
            price = int(rng.integers(-2**40, 2**40))
            disc, tax = int(rng.integers(-500, 500)), int(rng.integers(-500, 500))
        rows.append((int(rng.integers(ship_lo, ship_hi + 1)), qty, price, disc, tax,
                     RF[int(rng.integers(0, 3))], LS[int(rng.integers(0, 2))]))
    return rows


# ---- checks -------------------------------------------------------------------------
def check_todate(o) -> bool:
    if o.todate(1, 1, 1992) != FOR or o.todate(2, 9, 1998) != CUTOFF:
        return False
    if o.todate(1, 12, 1998) - 90 != CUTOFF:   # cpu/common.cpp:59-60
        return False
    d = datetime.date(1600, 1, 1)
    while d.year < 2101:
        if o.todate(d.day, d.month, d.year) != d.toordinal() + 365:
            return False
        d += datetime.timedelta(days=37)
    for y, m, dd in [(1996, 2, 29), (2000, 2, 29), (1900, 3, 1), (1999, 12, 31), (2000, 1, 1)]:
        if o.todate(dd, m, y) != datetime.date(y, m, dd).toordinal() + 365:
            return False
    return True


def _same(groups, expect) -> bool:
    return all(groups[g] == expect[g] for g in range(6))


def check_brute_force(o) -> bool:
    rng = np.random.default_rng(7)
    for n, tpch in [(1, True), (7, True), (64, True), (200, True), (50, False), (300, False)]:
        rows = random_rows(rng, n, tpch=tpch)
        cols = ref_columns(rows)
        for cut in (CUTOFF, 729906, 730088):
            got, unexpected = o.naive(cols, 0, n, cut)
            if unexpected or not _same(got, py_q1(rows, cut)):
                return False
    return True


def check_closed_form(o) -> bool:
    # n identical rows in every group: each sum is n times the row's value
    n = 1000
    qty, price, disc, tax = 1700, 2345678, 4, 7
    for rf in RF:
        for ls in LS:
            rows = [(CUTOFF, qty, price, disc, tax, rf, ls)] * n
            got, _ = o.naive(ref_columns(rows), 0, n, CUTOFF)
            g = got[slot(rf, ls)]
            want = dict(sum_qty=n * qty, sum_base_price=n * price, sum_disc_price=n * (100 - disc) * price,
                        sum_charge=n * (100 - disc) * price * (100 + tax), sum_disc=n * disc, count=n)
            if g != want or any(got[s]["count"] for s in range(6) if s != slot(rf, ls)):
                return False
    return True


def check_filter_boundary(o) -> bool:
    rows = [(CUTOFF - 1, 100, 1000, 1, 1, "A", "F"), (CUTOFF, 200, 2000, 2, 2, "A", "F"),
            (CUTOFF + 1, 400, 4000, 3, 3, "A", "F")]
    got, _ = o.naive(ref_columns(rows), 0, 3, CUTOFF)
    return got[0]["count"] == 2 and got[0]["sum_qty"] == 300 and got[0]["sum_base_price"] == 3000


def check_int128(o) -> bool:
    # per-row charge near 9e18 (fits int64); 5 rows pass 2^63 and 2^64; negative too
    rows = [(CUTOFF, 1, 9 * 10**14, 0, 0, "N", "O")] * 5 + [(CUTOFF, 1, 9 * 10**14, 200, 0, "R", "F")] * 5
    cols = ref_columns(rows)
    got, _ = o.naive(cols, 0, len(rows), CUTOFF)
    return _same(got, py_q1(rows, CUTOFF)) and got[slot("N", "O")]["sum_charge"] > 2**64 and \
        got[slot("R", "F")]["sum_charge"] < -2**64


def check_hyper(o) -> bool:
    rng = np.random.default_rng(11)
    rows = random_rows(rng, 300, tpch=True)
    cols = ref_columns(rows)
    h, flag = o.hyper(cols, 0, len(rows), CUTOFF)
    n, _ = o.naive(cols, 0, len(rows), CUTOFF)
    if flag or h != n:
        return False
    over = ref_columns([(CUTOFF, 1, 2**62, 0, 0, "A", "F")])  # (100 - 0) * 2^62 overflows int64
    _, flag = o.hyper(over, 0, 1, CUTOFF)
    if flag != 1:
        return False
    # morsel-parallel: the same sums, and an overflow inside any morsel (first, middle,
    # last row) or only in the merge of the workers' tables raises the flag
    big = random_rows(rng, 3001, tpch=True)
    bcols = ref_columns(big)
    want, _ = o.naive(bcols, 0, 3001, CUTOFF)
    for threads, morsel in ((1, 100), (4, 77), (5, 1000)):
        got, flag = o.hyper_parallel(bcols, 0, 3001, CUTOFF, threads, morsel)
        if flag or got != want:
            return False
    for at in (0, 1500, 3000):
        rows = list(big)
        rows[at] = (CUTOFF, 1, 2**62, 0, 0, "A", "F")
        if o.hyper_parallel(ref_columns(rows), 0, 3001, CUTOFF, 3, 100)[1] != 1:
            return False
    # one row fits, two rows' quantity sum passes 2^63 (whether one worker adds both rows or
    # the merge adds two workers' tables)
    near = ref_columns([(CUTOFF, 2**62, 1, 0, 0, "A", "F")] * 2)
    _, f1 = o.hyper(near, 0, 1, CUTOFF)
    _, f2 = o.hyper_parallel(near, 0, 2, CUTOFF, 2, 1)
    return f1 == 0 and f2 == 1


def _pack(codes, bits):
    per = 32 // bits
    words = np.zeros((len(codes) + per - 1) // per, np.uint64)
    for i, c in enumerate(codes):
        words[i // per] |= np.uint64(int(c) << (bits * (i % per)))
    return words.astype(np.uint32)


def check_compress(o) -> bool:
    rng = np.random.default_rng(3)
    for n in (1, 15, 16, 17, 31, 32, 33, 100, 1000):
        rows = random_rows(rng, n, tpch=True)
        cols = ref_columns(rows)
        c = o.compress(cols, n)
        if not np.array_equal(c["ship_date"].astype(np.int64), cols["ship_date"].astype(np.int64) - FOR):
            return False
        if not np.array_equal(c["quantity"].astype(np.int64), cols["quantity"] // 100):
            return False
        for k in ("extended_price", "discount", "tax"):
            if not np.array_equal(c[k].astype(np.int64), cols[k]):
                return False
        rf_codes = [{"A": 0, "N": 1, "R": 2}[r[5]] for r in rows]
        ls_codes = [{"F": 0, "O": 1}[r[6]] for r in rows]
        if not np.array_equal(c["return_flag"], _pack(rf_codes, 2)):
            return False
        if not np.array_equal(c["line_status"], _pack(ls_codes, 1)):
            return False
        for i in {0, 15, 16, 31, 32, 33, n - 1} & set(range(n)):
            if o.get_return_flag(c["return_flag"], i) != rf_codes[i] or \
                    o.get_line_status(c["line_status"], i) != ls_codes[i]:
                return False
    return True


def check_compressed_kernel(o) -> bool:
    rng = np.random.default_rng(5)
    n = 2000
    rows = random_rows(rng, n, tpch=True)
    cols = ref_columns(rows)
    c = o.compress(cols, n)
    for cut in (CUTOFF, 729906, 730088, FOR - 5, FOR + 65536):
        want = py_q1(rows, cut)
        got = o.compressed_kernel(c, 0, n, cut)
        for g in got:
            g["sum_qty"] *= 100
        if not _same(got, want):
            return False
        # sub-ranges starting mid-word
        for b, e in ((3, 977), (17, 1999), (32, 64)):
            sub = o.compressed_kernel(c, b, e, cut)
            for g in sub:
                g["sum_qty"] *= 100
            if not _same(sub, py_q1(rows[b:e], cut)):
                return False
    return True


def check_compressed_parallel(o) -> bool:
    # morsel-parallel compressed kernel vs the independent formulation: morsels and
    # sub-ranges that start and end mid-word, more threads than morsels
    rng = np.random.default_rng(29)
    n = 3001
    rows = random_rows(rng, n, tpch=True)
    c = o.compress(ref_columns(rows), n)
    for cut in (CUTOFF, 729906):
        for b, e in ((0, n), (5, 2999), (17, 40)):
            want = py_q1(rows[b:e], cut)
            for threads, morsel in ((1, 100), (3, 77), (4, 1000), (7, 33)):
                got = o.compressed_parallel(c, b, e, cut, threads, morsel)
                for g in got:
                    g["sum_qty"] *= 100
                if not _same(got, want):
                    return False
    return True


def check_filter_bits(o) -> bool:
    rng = np.random.default_rng(9)
    for n in (1, 31, 32, 33, 95, 1000):
        ship = rng.integers(0, 2600, n).astype(np.uint16)
        for cut in (CUTOFF, 729906):
            bits = o.precompute_filter(ship, n, cut)
            want = _pack((ship.astype(np.int64) <= cut - FOR).astype(int), 1)
            if not np.array_equal(bits, want):
                return False
    return True


def check_pushdown(o) -> bool:
    rng = np.random.default_rng(13)
    n = 1500
    rows = random_rows(rng, n, tpch=True)
    c = o.compress(ref_columns(rows), n)
    for cut in (CUTOFF, 729906):
        f = o.precompute_filter(c["ship_date"], n, cut)
        if o.pushdown_kernel(f, c, 0, n) != o.compressed_kernel(c, 0, n, cut):
            return False
    return True


def check_cutoff_shift(o) -> bool:
    # The reference fixes the cutoff; a run at cutoff c equals the fixed-cutoff run on
    # ship dates shifted by (CUTOFF - c), because only the filter reads the ship date.
    rng = np.random.default_rng(17)
    rows = random_rows(rng, 500, tpch=True)
    cols = ref_columns(rows)
    for c in (729906, 730088, 728000):
        shifted = dict(cols)
        shifted["ship_date"] = (cols["ship_date"] + (CUTOFF - c)).astype(np.int32)
        if o.naive(cols, 0, 500, c) != o.naive(shifted, 0, 500, CUTOFF):
            return False
    return True


def check_invariants(o) -> bool:
    rng = np.random.default_rng(19)
    n = 20000
    rows = random_rows(rng, n, tpch=True)
    cols = ref_columns(rows)
    got, _ = o.naive(cols, 0, n, CUTOFF)
    m = cols["ship_date"] <= CUTOFF
    if sum(g["count"] for g in got) != int(m.sum()):
        return False
    if sum(g["sum_qty"] for g in got) != int(cols["quantity"][m].sum()):
        return False
    if sum(g["sum_disc"] for g in got) != int(cols["discount"][m].sum()):
        return False
    p = cols["extended_price"][m].astype(object)
    d = cols["discount"][m].astype(object)
    t = cols["tax"][m].astype(object)
    return sum(g["sum_charge"] for g in got) == int(((100 - d) * p * (100 + t)).sum())


def check_parallel(o) -> bool:
    rng = np.random.default_rng(23)
    rows = random_rows(rng, 5000, tpch=True)
    cols = ref_columns(rows)
    return o.naive_parallel(cols, 0, 5000, CUTOFF, 4, 100) == o.naive(cols, 0, 5000, CUTOFF)


def check_x100(o) -> bool:
    """KernelX100 over ComprData equals the independent formulation except for the
    reference's sum_disc quirk: it sums disc_1 = 100 - disc, i.e. 100 * count - sum(disc).
    Sizes and morsels leave partial 1024-row vectors inside and at the end of morsels."""
    rng = np.random.default_rng(29)
    for n in (1, 1023, 1024, 1025, 5003):
        rows = random_rows(rng, n, tpch=True)
        cols = ref_columns(rows)
        comp = o.compact(cols, n)
        for cut in (CUTOFF, 729906, 730088):
            want = py_q1(rows, cut)
            for g in want:
                g["sum_disc"] = 100 * g["count"] - g["sum_disc"]
            for threads, morsel in ((1, 10 * 1024), (3, 1500)):
                got, unexpected = o.x100(comp, 0, n, cut, threads, morsel)
                if unexpected or not _same(got, want):
                    return False
    return True


def check_avx512(o) -> bool:
    """The AVX-512 kernel's 8-lane table equals the independent formulation exactly while
    every lane's sums stay below 2^63 (all TPC-H-domain inputs); sizes that are not
    multiples of 16 exercise the cardinality mask."""
    rng = np.random.default_rng(31)
    for n in (1, 7, 15, 16, 17, 200, 1003):
        rows = random_rows(rng, n, tpch=True)
        cols = ref_columns(rows)
        for cut in (CUTOFF, 729906):
            got, unexpected = o.avx512(cols, 0, n, cut)
            if unexpected or not _same(got, py_q1(rows, cut)):
                return False
    return True


CHECKS = {name[len("check_"):]: fn for name, fn in list(globals().items()) if name.startswith("check_")}


def run_all(o) -> dict:
    return {name: bool(fn(o)) for name, fn in CHECKS.items()}
