"""Python handle on the continuous-batching server of the C ABI (q1_server_*), plus the
open-loop load generator used for the serving measurement (Poisson arrivals over a set of
HBM-resident tables, random cutoffs and row ranges)."""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from . import q1


class Server:
    def __init__(self, ctx: q1.Context, max_batch=256, max_inflight=2, queue_capacity=65536,
                 batch_window_us=0, memory_budget=0, exclusive_bytes=0):
        """memory_budget: device bytes the in-flight work may pin (the columns of each table
        with requests in flight, plus descriptor and result state per request); 0 = three
        quarters of the device's memory.  exclusive_bytes: requests that scan at least this
        many column bytes run alone and one at a time; 0 = 512 MiB, 2**64 - 1 = never."""
        if not hasattr(q1.lib, "q1_server_create"):
            raise RuntimeError("libq1hip.so was built without the server")
        self.ctx = ctx
        cfg = q1.ServerConfig(max_batch, max_inflight, queue_capacity, batch_window_us, memory_budget,
                              exclusive_bytes)
        h = C.c_void_p()
        ctx.check(q1.lib.q1_server_create(ctx.handle, C.byref(cfg), C.byref(h)), "q1_server_create")
        self.handle = h.value

    def submit(self, table: q1.Table, row_begin: int, row_count: int, cutoff_day: int) -> tuple[int, int]:
        """Returns (status, ticket)."""
        req = q1.Request(table.handle, row_begin, row_count, cutoff_day, 0)
        ticket = C.c_uint64()
        st = q1.lib.q1_server_submit(self.handle, C.byref(req), C.byref(ticket))
        return st, ticket.value

    def wait(self, ticket: int, timeout_us: int = -1):
        """Returns (status, groups or None, stats)."""
        r, s = q1.Result(), q1.RequestStats()
        st = q1.lib.q1_server_wait(self.handle, ticket, timeout_us, C.byref(r), C.byref(s))
        return st, (q1.result_to_list(r) if st == q1.OK else None), s

    def cancel(self, ticket: int) -> int:
        return q1.lib.q1_server_cancel(self.handle, ticket)

    def drain(self):
        self.ctx.check(q1.lib.q1_server_drain(self.handle), "q1_server_drain")

    def close(self):
        if self.handle:
            q1.lib.q1_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def make_request_stream(tables, rate_per_s: float, duration_s: float, seed: int = 4):
    """Poisson arrival times and random (table, row range, cutoff) per request."""
    rng = np.random.default_rng(seed)
    n = int(rate_per_s * duration_s)
    gaps = rng.exponential(1.0 / rate_per_s, n)
    t = np.cumsum(gaps)
    reqs = []
    for i in range(n):
        tab = tables[int(rng.integers(0, len(tables)))]
        size = int(rng.integers(1, tab.rows + 1))
        begin = int(rng.integers(0, tab.rows - size + 1))
        cutoff = 729906 + int(rng.integers(0, 183))  # 1998-06-01 .. 1998-11-30
        reqs.append((float(t[i]), tab, begin, size, cutoff))
    return reqs


def run_open_loop(server: Server, stream, collect=None):
    """Submits every request at its arrival time from one thread while another waits for
    completions; returns per-request (latency_us, batch_size) and the wall time."""
    tickets = [None] * len(stream)
    done = [None] * len(stream)
    submitted = threading.Semaphore(0)

    def waiter():
        for i in range(len(stream)):
            submitted.acquire()
            st, groups, s = server.wait(tickets[i])
            done[i] = (st, s.latency_us, s.batch_size)
            if collect is not None:
                collect(i, st, groups)

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for i, (at, tab, b, n, c) in enumerate(stream):
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        while True:
            st, tk = server.submit(tab, b, n, c)
            if st != q1.EBUSY:
                break
            time.sleep(50e-6)
        if st != q1.OK:
            raise q1.Q1Error(st, "q1_server_submit")
        tickets[i] = tk
        submitted.release()
    th.join()
    wall = time.perf_counter() - t0
    return done, wall
