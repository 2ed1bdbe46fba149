"""HIP path vs the oracle, bit-exact, through the C ABI (every layout x placement, edge
sizes, row offsets, runtime cutoffs, launch shapes, out-of-domain values, 128-bit sums)."""
import sys

import numpy as np
import pytest

import pinning
from tpchq01_gpu_amd import q1

pytestmark = pytest.mark.gpu

SEED = 0x51C00001
LAYOUTS = [q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED, q1.COMPRESSED_PUSHDOWN]
PLACEMENTS = [q1.IN_REGISTERS, q1.IN_REGISTERS_PER_THREAD, q1.LOCAL_MEM, q1.PER_THREAD_SHARED_MEM, q1.GLOBAL]
SIZES = [0, 1, 15, 16, 17, 31, 32, 33, 1000, 65_537, 1_000_003]
CUTOFFS = [729999, 729906, 730088]   # 1998-09-02, 1998-06-01, 1998-11-30


def as_dicts(groups):
    return [dict(sum_qty=g.sum_qty, sum_base_price=g.sum_base_price, sum_disc_price=g.sum_disc_price,
                 sum_charge=g.sum_charge, sum_disc=g.sum_disc, count=g.count) for g in groups]


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_host_cache = {}


def host_tables(rows):
    if rows not in _host_cache:
        _host_cache.clear()
        _host_cache[rows] = {
            q1.REFERENCE: q1.generate_host(q1.REFERENCE, SEED, rows),
            q1.UNCOMPRESSED: q1.generate_host(q1.UNCOMPRESSED, SEED, rows),
            q1.COMPRESSED: q1.generate_host(q1.COMPRESSED, SEED, rows),
        }
    return _host_cache[rows]


def device_table(ctx, layout, rows, cutoff):
    h = host_tables(rows)
    if layout == q1.COMPRESSED_PUSHDOWN:
        comp = ctx.upload(q1.COMPRESSED, h[q1.COMPRESSED], rows)
        return ctx.build_filter(comp, cutoff)
    return ctx.upload(layout, h[layout], rows)


@pytest.mark.parametrize("rows", SIZES)
def test_all_layouts_and_placements(ctx, oracle, rows):
    ref = host_tables(rows)[q1.REFERENCE]
    for cutoff in CUTOFFS:
        want, _ = oracle.naive(ref, 0, rows, cutoff)
        for layout in LAYOUTS:
            t = device_table(ctx, layout, rows, cutoff)
            for p in PLACEMENTS:
                note(f"case rows={rows} cutoff={cutoff} layout={q1.LAYOUT_NAMES[layout]} placement={q1.PLACEMENT_NAMES[p]}")
                got = ctx.run(t, 0, rows, cutoff, p)
                assert as_dicts(got) == want
            t.free()


@pytest.mark.parametrize("layout", LAYOUTS)
def test_row_ranges(ctx, oracle, layout):
    rows = 100_000
    ref = host_tables(rows)[q1.REFERENCE]
    cutoff = 729906
    t = device_table(ctx, layout, rows, cutoff)
    for begin, end in [(3, rows - 5), (17, rows), (0, rows - 1), (31, 33), (15, 16), (4097, 70001), (5, 5)]:
        want, _ = oracle.naive(ref, begin, end, cutoff)
        for p in PLACEMENTS:
            note(f"case layout={q1.LAYOUT_NAMES[layout]} range=[{begin},{end}) placement={p}")
            assert as_dicts(ctx.run(t, begin, end - begin, cutoff, p)) == want


@pytest.mark.parametrize("layout", LAYOUTS)
def test_range_starts_around_the_128_row_alignment(ctx, oracle, layout):
    """A launch's blocks start at its first row rounded down to 128 rows and mask the rows
    before it: starts on each side of 128-row boundaries, ranges that end inside the first
    aligned stretch, and launch shapes whose blocks hold fewer rows than that stretch."""
    rows = 70_000
    ref = host_tables(rows)[q1.REFERENCE]
    cutoff = 729999
    t = device_table(ctx, layout, rows, cutoff)
    ranges = [(1, rows - 1), (64, rows - 64), (127, 129), (128, 256), (129, 300), (255, 1), (100, 10), (5, 50),
              (256 * 100 + 1, 40_000), (rows - 129, 128), (rows - 1, 1)]
    for begin, n in ranges:
        want, _ = oracle.naive(ref, begin, begin + n, cutoff)
        for p in PLACEMENTS:
            for shape in (None, q1.make_launch(64, 16)):
                note(f"case layout={q1.LAYOUT_NAMES[layout]} range=[{begin},{begin + n}) placement={p} "
                     f"shape={'tuned' if shape is None else '64x16'}")
                assert as_dicts(ctx.run(t, begin, n, cutoff, p, shape)) == want
    t.free()


def tuned_shapes():
    """Every distinct (layout, placement, threads, tuples) of the shipped tuning table's
    "launch" and "launch_at" lines."""
    import os
    from tpchq01_gpu_amd import PACKAGE_DIR
    shapes = set()
    for line in open(os.path.join(PACKAGE_DIR, "tuning", "gfx950.txt")):
        f = line.split()
        if f and f[0] == "launch":
            shapes.add(tuple(map(int, f[1:5])))
        elif f and f[0] == "launch_at":
            shapes.add(tuple(map(int, f[2:6])))
    return sorted(shapes)


def test_every_tuned_shape(ctx, oracle):
    """Each launch shape the tuning table can hand out, for its own layout and placement,
    on a 1M-row range starting off a tile boundary: equal to the oracle."""
    rows, b, e = 1_000_003, 37, 1_000_003
    ref = host_tables(rows)[q1.REFERENCE]
    want, _ = oracle.naive(ref, b, e, 729999)
    shapes = tuned_shapes()
    assert len(shapes) >= 20
    for layout in LAYOUTS:
        t = device_table(ctx, layout, rows, 729999)
        for l, p, tpb, tpt in shapes:
            if l != layout:
                continue
            note(f"case tuned shape layout={q1.LAYOUT_NAMES[l]} placement={q1.PLACEMENT_NAMES[p]} {tpb}x{tpt}")
            got = ctx.run(t, b, e - b, 729999, p, q1.make_launch(tpb, tpt))
            assert as_dicts(got) == want, (l, p, tpb, tpt)
        t.free()


@pytest.mark.parametrize("shape", [(64, 16), (128, 256), (128, 48), (256, 4096), (256, 16384), (64, 1024)])
def test_launch_shapes(ctx, oracle, shape):
    rows = 300_001
    ref = host_tables(rows)[q1.REFERENCE]
    want, _ = oracle.naive(ref, 0, rows, 729999)
    for layout in (q1.COMPRESSED, q1.UNCOMPRESSED, q1.REFERENCE):
        t = device_table(ctx, layout, rows, 729999)
        for p in PLACEMENTS:
            note(f"case shape={shape} layout={layout} placement={p}")
            got = ctx.run(t, 0, rows, 729999, p, q1.make_launch(*shape))
            assert as_dicts(got) == want
        t.free()


@pytest.mark.parametrize("layout", LAYOUTS)
def test_spread_merge_large_grids(ctx, oracle, layout):
    """Grids of kSpreadMinBlocks (176) blocks or more merge through the context's spread
    tables and a fold kernel; launches one after another must each see clean tables."""
    rows = 2_000_003
    ref = host_tables(rows)[q1.REFERENCE]
    cutoff = 729906
    want, _ = oracle.naive(ref, 0, rows, cutoff)
    t = device_table(ctx, layout, rows, cutoff)
    tile = {q1.REFERENCE: 4, q1.UNCOMPRESSED: 8}.get(layout, 16)
    for p in PLACEMENTS:
        for tpb, tpt in ((64, 16), (256, 16), (64, 32), (128, 64)):
            blocks = -(-rows // (tpb * tpt))
            note(f"case layout={q1.LAYOUT_NAMES[layout]} placement={p} shape={tpb}x{tpt} blocks={blocks}")
            for rep in range(2):
                got = ctx.run(t, 0, rows, cutoff, p, q1.make_launch(tpb, tpt))
                assert as_dicts(got) == want, (p, tpb, tpt, rep)
        # an offset range whose first and last tiles are partial
        b, e = 3 * tile + 1, rows - 5
        want_r, _ = oracle.naive(ref, b, e, cutoff)
        assert as_dicts(ctx.run(t, b, e - b, cutoff, p, q1.make_launch(64, 16))) == want_r
    t.free()


SPREAD_MIN_BLOCKS = 176   # kSpreadMinBlocks of the library


def test_spread_merge_threshold(ctx, oracle):
    """Grids of 175 blocks merge directly, 176 and more through the spread tables: same
    answers (64 lanes x 16 rows = 1024 rows per block)."""
    k = SPREAD_MIN_BLOCKS
    for rows in (1, (k - 1) * 1024, (k - 1) * 1024 + 1, k * 1024, 1023 * 1024 + 1):
        ref = q1.generate_host(q1.REFERENCE, SEED + 3, rows)
        comp = q1.generate_host(q1.COMPRESSED, SEED + 3, rows)
        want, _ = oracle.naive(ref, 0, rows, 729999)
        t = ctx.upload(q1.COMPRESSED, comp, rows)
        for p in (q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM, q1.GLOBAL):
            note(f"case rows={rows} placement={p}")
            assert as_dicts(ctx.run(t, 0, rows, 729999, p, q1.make_launch(64, 16))) == want
        t.free()


def test_cutoff_extremes(ctx, oracle):
    rows = 50_000
    ref = host_tables(rows)[q1.REFERENCE]
    t = device_table(ctx, q1.COMPRESSED, rows, 729999)
    for cutoff in (0, 727563 - 2, 727563, 727564, 730089, 727563 + 65535, 727563 + 70000, 2**31 - 1):
        want, _ = oracle.naive(ref, 0, rows, cutoff)
        for p in PLACEMENTS:
            assert as_dicts(ctx.run(t, 0, rows, cutoff, p)) == want, (cutoff, p)


def test_device_generator_matches_host(ctx):
    for layout in (q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED):
        for rows in (1, 33, 1000, 100_017):
            t = ctx.generate(layout, SEED + 9, rows)
            got = t.download()
            want = q1.generate_host(layout, SEED + 9, rows)
            for k in want:
                assert np.array_equal(got[k], want[k]), (layout, rows, k)
            t.free()


def test_device_filter_matches_oracle(ctx, oracle):
    rows = 77_777
    comp = host_tables(rows)[q1.COMPRESSED]
    t = ctx.upload(q1.COMPRESSED, comp, rows)
    for cutoff in CUTOFFS + [727000, 740000]:
        pd = ctx.build_filter(t, cutoff)
        got = pd.download()["filter"]
        assert np.array_equal(got, oracle.precompute_filter(comp["ship_date"], rows, cutoff))
    # sizes around the producer's 8-row lanes, 32-row words and 8,192-row blocks
    big = q1.generate_host(q1.COMPRESSED, SEED + 11, 40 * 16384)
    for n in (1, 7, 8, 9, 31, 32, 33, 255, 257, 2049, 8191, 8192, 8193, 3 * 8192 + 9, 16385,
              40 * 16384 - 24):
        c = {k: v[: (n + 15) // 16 if k == "return_flag" else (n + 31) // 32 if k == "line_status" else n].copy()
             for k, v in big.items()}
        tn = ctx.upload(q1.COMPRESSED, c, n)
        for cutoff in (729999, 729906):
            pd = ctx.build_filter(tn, cutoff)
            assert np.array_equal(pd.download()["filter"], oracle.precompute_filter(c["ship_date"], n, cutoff)), n
            pd.free()
        tn.free()


def test_uploaded_pushdown_columns(ctx, oracle):
    rows = 65_536 + 77
    ref = host_tables(rows)[q1.REFERENCE]
    comp = host_tables(rows)[q1.COMPRESSED]
    pdc = q1.pushdown_columns(comp, 729906, rows)
    t = ctx.upload(q1.COMPRESSED_PUSHDOWN, pdc, rows)
    want, _ = oracle.naive(ref, 0, rows, 729906)
    for p in PLACEMENTS:
        # the pushdown table's bitmap holds the filter; the cutoff argument is not used
        assert as_dicts(ctx.run(t, 0, rows, 0, p)) == want


def test_out_of_domain_values_take_the_64bit_path(ctx, oracle):
    rng = np.random.default_rng(5)
    rows = pinning.random_rows(rng, 20_000, tpch=False)
    cols = pinning.ref_columns(rows)
    want, _ = oracle.naive(cols, 0, len(rows), 729999)
    t = ctx.upload(q1.REFERENCE, cols, len(rows))
    for p in PLACEMENTS:
        assert as_dicts(ctx.run(t, 0, len(rows), 729999, p)) == want
    assert ctx.last_timing().placement == q1.IN_REGISTERS_PER_THREAD


def test_compressed_out_of_domain(ctx, oracle):
    # discount above 100 makes (100 - disc) negative: the 64-bit kernel must take over
    rows = 10_000
    comp = {k: v.copy() for k, v in host_tables(rows)[q1.COMPRESSED].items()}
    ref = {k: v.copy() for k, v in host_tables(rows)[q1.REFERENCE].items()}
    comp["discount"][::7] = 200
    ref["discount"][::7] = 200
    comp["extended_price"][::11] = 4_000_000_000
    ref["extended_price"][::11] = 4_000_000_000
    want, _ = oracle.naive(ref, 0, rows, 729999)
    t = ctx.upload(q1.COMPRESSED, comp, rows)
    for p in PLACEMENTS:
        assert as_dicts(ctx.run(t, 0, rows, 729999, p)) == want


@pytest.mark.parametrize("layout", [q1.REFERENCE, q1.UNCOMPRESSED, q1.COMPRESSED])
def test_mixed_domain_tiles(ctx, oracle, layout):
    """A table outside the TPC-H domain only in scattered rows: the any-value kernel takes
    the narrow path for in-domain tiles and the 64-bit path for the others, within the
    same wavefronts.  Every placement, several launch shapes, exact vs the oracle."""
    rows = 300_007
    ref = {k: v.copy() for k, v in host_tables(rows)[q1.REFERENCE].items()}
    rng = np.random.default_rng(11)
    bad = rng.choice(rows, 40, replace=False)
    if layout == q1.COMPRESSED:
        # compressed columns hold u8 discount/tax/quantity and u32 price
        ref["discount"][bad[:10]] = 200
        ref["tax"][bad[10:20]] = 99
        ref["extended_price"][bad[20:30]] = 3_000_000_000
        ref["quantity"][bad[30:]] = 9900        # 99 units
        dev = q1.compress_columns(ref, rows)
    else:
        ref["discount"][bad[:10]] = -5
        ref["tax"][bad[10:20]] = 1 << 40
        ref["extended_price"][bad[20:30]] = 1 << 45
        ref["quantity"][bad[30:]] = 6301
        if layout == q1.UNCOMPRESSED:
            # int32 quantity, int8 discount/tax: keep values the layout can hold
            ref["tax"][bad[10:20]] = 100
            dev = {"ship_date": ref["ship_date"], "quantity": ref["quantity"].astype(np.int32),
                   "extended_price": ref["extended_price"], "discount": ref["discount"].astype(np.int8),
                   "tax": ref["tax"].astype(np.int8), "return_flag": ref["return_flag"],
                   "line_status": ref["line_status"]}
        else:
            dev = ref
    want, _ = oracle.naive(ref, 0, rows, 729999)
    t = ctx.upload(layout, dev, rows)
    for p in PLACEMENTS:
        for shape in (None, (64, 16), (256, 1024)):
            note(f"case layout={q1.LAYOUT_NAMES[layout]} placement={p} shape={shape}")
            got = ctx.run(t, 0, rows, 729999, p, q1.make_launch(*shape) if shape else None)
            assert as_dicts(got) == want, (p, shape)
    assert ctx.last_timing().placement == q1.IN_REGISTERS_PER_THREAD   # the 64-bit kernel ran
    t.free()


def test_int128_sums_cross_2_64_general_path(ctx, oracle):
    rows = [(729999, 1, 9 * 10**14, 0, 0, "N", "O")] * 4000 + [(729999, 1, 9 * 10**14, 200, 0, "R", "F")] * 4000
    cols = pinning.ref_columns(rows)
    want, _ = oracle.naive(cols, 0, len(rows), 729999)
    assert want[1 << 1 | 1]["sum_charge"] > 2**64 and want[2 << 1]["sum_charge"] < -(2**64)
    t = ctx.upload(q1.REFERENCE, cols, len(rows))
    for p in PLACEMENTS:
        assert as_dicts(ctx.run(t, 0, len(rows), 729999, p)) == want


@pytest.mark.slow
def test_int128_carry_fast_path(ctx):
    """100M rows of the largest TPC-H-domain values in one group: sum_charge passes 2^64,
    so the carry of the per-block 128-bit merge is exercised.  Closed form."""
    n = 100_000_000
    price, disc, tax, qty = (1 << 24) - 1, 0, 15, 63
    cols = {
        "ship_date": np.zeros(n, np.uint16), "quantity": np.full(n, qty, np.uint8),
        "extended_price": np.full(n, price, np.uint32), "discount": np.full(n, disc, np.uint8),
        "tax": np.full(n, tax, np.uint8),
        "return_flag": np.full((n + 15) // 16, 0x55555555, np.uint32),     # every row 'N'
        "line_status": np.full((n + 31) // 32, 0xFFFFFFFF, np.uint32),     # every row 'O'
    }
    t = ctx.upload(q1.COMPRESSED, cols, n)
    charge = n * (100 - disc) * price * (100 + tax)
    assert charge > 2**64
    # the tuned shapes and 64 x 64 (24415 blocks) merge through the spread tables
    for p, launch in ((q1.IN_REGISTERS, None), (q1.PER_THREAD_SHARED_MEM, None), (q1.IN_REGISTERS_PER_THREAD, None),
                      (q1.IN_REGISTERS, q1.make_launch(64, 64)), (q1.PER_THREAD_SHARED_MEM, q1.make_launch(64, 64))):
        g = ctx.run(t, 0, n, 729999, p, launch)[1 << 1 | 1]
        assert (g.count, g.sum_qty, g.sum_base_price, g.sum_disc) == (n, n * qty * 100, n * price, 0)
        assert g.sum_disc_price == n * (100 - disc) * price
        assert g.sum_charge == charge
    t.free()


def test_shared_narrow_price_bound(ctx):
    """At 256 rows per lane the shared placement keeps a lane's price sum in 32 bits:
    256 rows of the largest domain price fill it to 2^32 - 256.  272 rows per lane take
    the 64-bit price field.  Closed form."""
    price = (1 << 24) - 1
    for tpt in (256, 272):
        n = 64 * tpt * 5 + 7          # five full blocks of 64 lanes and a partial one
        cols = {
            "ship_date": np.zeros(n, np.uint16), "quantity": np.full(n, 50, np.uint8),
            "extended_price": np.full(n, price, np.uint32), "discount": np.full(n, 10, np.uint8),
            "tax": np.full(n, 8, np.uint8),
            "return_flag": np.full((n + 15) // 16, 0xAAAAAAAA, np.uint32),   # every row 'R'
            "line_status": np.zeros((n + 31) // 32, np.uint32),              # every row 'F'
        }
        t = ctx.upload(q1.COMPRESSED, cols, n)
        g = ctx.run(t, 0, n, 729999, q1.PER_THREAD_SHARED_MEM, q1.make_launch(64, tpt))[2 << 1]
        assert (g.count, g.sum_qty, g.sum_base_price, g.sum_disc) == (n, n * 5000, n * price, n * 10), tpt
        assert g.sum_disc_price == n * 90 * price and g.sum_charge == n * 90 * 108 * price, tpt
        t.free()


def test_split_matches_oracle(ctx, oracle):
    rows = 1_000_003
    h = host_tables(rows)
    want, _ = oracle.naive(h[q1.REFERENCE], 0, rows, 729999)
    for layout in (q1.COMPRESSED, q1.REFERENCE):
        t = ctx.upload(layout, h[layout], rows)
        for frac in (0.0, 0.25, 0.5, 1.0):
            got, cms, gms = ctx.run_split(t, h[layout], rows, frac, 729999, q1.IN_REGISTERS,
                                          align_rows=1 << 10, cpu_threads=4)
            assert as_dicts(got) == want, (layout, frac)
        t.free()


def test_streaming_matches_oracle(ctx, oracle):
    rows = 3_000_017
    h = host_tables(rows)
    want, _ = oracle.naive(h[q1.REFERENCE], 0, rows, 729906)
    for layout in (q1.COMPRESSED, q1.UNCOMPRESSED):
        for streams, chunk in ((1, 1 << 20), (4, 1 << 18), (3, 96_000)):
            for p in (q1.IN_REGISTERS, q1.PER_THREAD_SHARED_MEM, q1.GLOBAL):
                got = ctx.run_streaming(layout, h[layout], rows, 729906, p,
                                        q1.make_launch(flags=q1.LAUNCH_ASSUME_TPCH_DOMAIN), chunk, streams)
                assert as_dicts(got) == want, (layout, streams, chunk, p)
    # without the domain flag the 64-bit kernels run, same answer
    got = ctx.run_streaming(q1.COMPRESSED, h[q1.COMPRESSED], rows, 729906, q1.IN_REGISTERS, None, 1 << 20, 2)
    assert as_dicts(got) == want


@pytest.mark.parametrize("pattern", ["every_7th", "one_stretch"])
def test_unused_return_flag_code_never_counts(ctx, oracle, pattern):
    """The compressed return-flag column has a fourth 2-bit code (3) that no real row has;
    rows carrying it fall outside the six groups and are skipped.  Rows with it spread over
    the table (most tile groups hold one) or in one 64-row stretch (most tile groups hold
    none, so they take the kernels' path for groups known to be clean): every placement
    matches the oracle's compressed kernel."""
    n = 300_017
    clean = q1.generate_host(q1.COMPRESSED, SEED + 9, n)
    cols = {k: v.copy() for k, v in clean.items()}
    rows = np.arange(n)
    bad = rows[rows % 7 == 3] if pattern == "every_7th" else rows[100_000:100_064]
    for i in bad:
        cols["return_flag"][i // 16] |= np.uint32(3 << (2 * (i % 16)))
    t = ctx.upload(q1.COMPRESSED, cols, n)
    # 727563 + 65535: a compressed cutoff at which every ship date passes, so the kernels
    # cannot turn the planted rows into filtered ones and test each row's slot instead
    for cutoff in (729999, 727563 + 65535):
        want = [dict(g, sum_qty=g["sum_qty"] * 100) for g in oracle.compressed_kernel(cols, 0, n, cutoff)]
        # the planted rows passed the filter and no longer count
        assert sum(g["count"] for g in want) < sum(g["count"] for g in oracle.compressed_kernel(clean, 0, n, cutoff))
        for p in PLACEMENTS:
            for launch in (None, q1.make_launch(64, 16), q1.make_launch(256, 256)):
                note(f"case invalid rf {pattern} cutoff={cutoff} placement={p} "
                     f"launch={launch and (launch.threads_per_block, launch.tuples_per_thread)}")
                assert as_dicts(ctx.run(t, 0, n, cutoff, p, launch)) == want
    cutoff = 729999
    assert as_dicts(ctx.run(t, 5, n - 9, cutoff, q1.PER_THREAD_SHARED_MEM)) == \
        [dict(g, sum_qty=g["sum_qty"] * 100) for g in oracle.compressed_kernel(cols, 5, n - 4, cutoff)]
    t.free()


@pytest.mark.parametrize("tpt", [16, 128, 240, 256])
@pytest.mark.parametrize("mix", [False, True])
def test_shared_packed_slot_bounds(ctx, tpt, mix):
    """At most 256 rows per lane the LDS placement keeps a compressed lane's sums in three
    64-bit slots: price [0,32), discount [32,44) and quantity [44,58) in one; the discount
    product [0,36) and count [48,57) in another; the charge.  256 rows of the domain's
    largest values in one group fill each field to just below its top: price
    256 * (2^24 - 1) < 2^32, discount 256 * 15 < 2^12, quantity 256 * 63 < 2^14, product
    256 * 15 * (2^24 - 1) < 2^36.  Closed form, with partial blocks."""
    price = (1 << 24) - 1
    n = 256 * tpt * 3 + 77
    rows = np.arange(n)
    # every row in one group (each lane's slot of that group takes all of its rows), or mixed
    rf = (rows // 7) % 3 if mix else np.full(n, 2)
    ls = (rows // 5) % 2 if mix else np.zeros(n, np.int64)
    cols = {
        "ship_date": np.zeros(n, np.uint16), "quantity": np.full(n, 63, np.uint8),
        "extended_price": np.full(n, price, np.uint32), "discount": np.full(n, 15, np.uint8),
        "tax": np.full(n, 15, np.uint8),
        "return_flag": np.zeros((n + 15) // 16, np.uint32), "line_status": np.zeros((n + 31) // 32, np.uint32),
    }
    for i in range(n):
        cols["return_flag"][i // 16] |= np.uint32(int(rf[i]) << (2 * (i % 16)))
        cols["line_status"][i // 32] |= np.uint32(int(ls[i]) << (i % 32))
    t = ctx.upload(q1.COMPRESSED, cols, n)
    got = ctx.run(t, 0, n, 729999, q1.PER_THREAD_SHARED_MEM, q1.make_launch(256, tpt))
    for g in range(6):
        k = int(np.sum((rf << 1 | ls) == g))
        assert (got[g].count, got[g].sum_qty, got[g].sum_base_price, got[g].sum_disc) == (k, k * 6300, k * price, k * 15), g
        assert got[g].sum_disc_price == k * 85 * price and got[g].sum_charge == k * 85 * 115 * price, g
    t.free()
