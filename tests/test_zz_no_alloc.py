"""q1_enqueue_noalloc, the drop-in launch that allocates nothing (its scratch is a static
per-device array set up by q1_init), against q1_enqueue: the same device table after two
calls in a row, the second on a smaller input, for every placement, with grids that merge
through the spread tables (>= 176 blocks) and the global placement's per-block tables;
launches that need more scratch than the array holds are refused."""
import ctypes as C
import sys

import pytest

from tpchq01_gpu_amd import q1

pytestmark = pytest.mark.gpu

SEED = 0x51C00077


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(fn, ctx, p, layout, launch, cols, begin, n, cutoff, torch):
    agg = torch.zeros(48, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    ctx.check(fn(ctx.handle, p, layout, C.byref(launch) if launch else None, C.c_void_p(agg.data_ptr()),
                 C.byref(cols), begin, n, cutoff, None), fn.__name__)
    ctx.synchronize()
    return agg.cpu().numpy().copy()


def test_noalloc_matches_enqueue_on_two_calls_in_a_row():
    import torch
    with q1.Context(0) as ctx:
        rows = 3_000_017
        t = ctx.generate(q1.COMPRESSED, SEED, rows)
        cols = t.device_columns()
        cases = [
            (q1.PER_THREAD_SHARED_MEM, q1.make_launch(64, 16, 0, q1.LAUNCH_ASSUME_TPCH_DOMAIN)),   # spread merge
            (q1.IN_REGISTERS, q1.make_launch(64, 32, 0, q1.LAUNCH_ASSUME_TPCH_DOMAIN)),            # spread merge
            (q1.GLOBAL, q1.make_launch(128, 512, 0, q1.LAUNCH_ASSUME_TPCH_DOMAIN)),                 # per-block tables
            (q1.GLOBAL, q1.make_launch(64, 16, 0, q1.LAUNCH_ASSUME_TPCH_DOMAIN)),                   # both
            (q1.LOCAL_MEM, None),                                                                  # any-value kernel
            (q1.IN_REGISTERS_PER_THREAD, q1.make_launch(flags=q1.LAUNCH_ASSUME_TPCH_DOMAIN)),
        ]
        for p, launch in cases:
            # two calls in a row, the second on a smaller input
            for begin, n, cutoff in ((0, rows, 729999), (17, 1_000_003, 729906)):
                note(f"case noalloc placement={p} begin={begin} rows={n}")
                want = run(q1.lib.q1_enqueue, ctx, p, q1.COMPRESSED, launch, cols, begin, n, cutoff, torch)
                got = run(q1.lib.q1_enqueue_noalloc, ctx, p, q1.COMPRESSED, launch, cols, begin, n, cutoff, torch)
                assert (got == want).all(), (p, begin, n)
        # more per-block tables than the static array holds: refused, nothing launched
        big = q1.make_launch(64, 16, 40_000, q1.LAUNCH_ASSUME_TPCH_DOMAIN)
        agg = torch.zeros(48, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        st = q1.lib.q1_enqueue_noalloc(ctx.handle, q1.GLOBAL, q1.COMPRESSED, C.byref(big), C.c_void_p(agg.data_ptr()),
                                       C.byref(cols), 0, rows, 729999, None)
        assert st == q1.EINVAL
        assert b"static array" in q1.lib.q1_last_error(ctx.handle)
        t.free()
