"""ctypes binding of the C ABI in include/q1_hip.h (libq1hip.so, built in-tree).

This module mirrors the reference's host-side interface for the Q1 path: the
hash-table placements keyed like ``--hash-table-placement`` (src/main.cu:166-173 and the
README spellings), the three input modes (plain / compressed / filter pushdown,
src/main.cu:166-191), device-resident tables, and fully reduced results in the reference
CPU kernels' scales (cpu/common.hpp:158-165).  There is no fallback: if the HIP library
is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# Q1_LIB_PATH selects another build of the library (e.g. an experimental variant).
LIB_PATH = os.environ.get("Q1_LIB_PATH") or os.path.join(_HERE, "lib", "libq1hip.so")

NUM_GROUPS = 6

# q1_status
OK, EINVAL, ENOMEM, EDEVICE, EOVERFLOW, EBUSY, ETIMEOUT, ECANCELED = range(8)

# q1_placement
IN_REGISTERS, IN_REGISTERS_PER_THREAD, LOCAL_MEM, PER_THREAD_SHARED_MEM, GLOBAL, AUTO = range(6)
PLACEMENTS = {
    # code names (src/main.cu:166-173)
    "in_registers": IN_REGISTERS,
    "in_registers_per_thread": IN_REGISTERS_PER_THREAD,
    "local_mem": LOCAL_MEM,
    "shared_mem_per_thread": PER_THREAD_SHARED_MEM,
    "global": GLOBAL,
    # README spellings (README.md:53)
    "in-registers": IN_REGISTERS,
    "in-registers-per-thread": IN_REGISTERS_PER_THREAD,
    "local-mem": LOCAL_MEM,
    "per-thread-shared-mem": PER_THREAD_SHARED_MEM,
    "auto": AUTO,
}
PLACEMENT_NAMES = ["in_registers", "in_registers_per_thread", "local_mem", "shared_mem_per_thread", "global"]

# q1_layout
REFERENCE, UNCOMPRESSED, COMPRESSED, COMPRESSED_PUSHDOWN = range(4)
LAYOUT_NAMES = ["reference", "uncompressed", "compressed", "compressed_pushdown"]
BITS_PER_ROW = [304, 160, 75, 60]

LAUNCH_ASSUME_TPCH_DOMAIN = 1

COLUMN_NAMES = ["ship_date", "quantity", "extended_price", "discount", "tax",
                "return_flag", "line_status", "filter"]

# element dtype per layout and column (None = column unused); packed columns are uint32 words
_DTYPES = {
    REFERENCE: [np.int32, np.int64, np.int64, np.int64, np.int64, np.uint8, np.uint8, None],
    UNCOMPRESSED: [np.int32, np.int32, np.int64, np.int8, np.int8, np.uint8, np.uint8, None],
    COMPRESSED: [np.uint16, np.uint8, np.uint32, np.uint8, np.uint8, np.uint32, np.uint32, None],
    COMPRESSED_PUSHDOWN: [None, np.uint8, np.uint32, np.uint8, np.uint8, np.uint32, np.uint32, np.uint32],
}


class Q1Error(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what} (status {status})")
        self.status = status


class Columns(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in COLUMN_NAMES]


class Group(C.Structure):
    _fields_ = [("sum_qty", C.c_int64), ("sum_base_price", C.c_int64),
                ("sum_disc_price_lo", C.c_uint64), ("sum_disc_price_hi", C.c_int64),
                ("sum_charge_lo", C.c_uint64), ("sum_charge_hi", C.c_int64),
                ("sum_disc", C.c_int64), ("count", C.c_int64)]


class Result(C.Structure):
    _fields_ = [("g", Group * NUM_GROUPS)]


class DeviceAggregates(C.Structure):
    _fields_ = [(n, C.c_uint64 * NUM_GROUPS) for n in
                ["sum_qty", "sum_base_price", "sum_disc_price_lo", "sum_disc_price_hi",
                 "sum_charge_lo", "sum_charge_hi", "sum_disc", "count"]]


class Launch(C.Structure):
    _fields_ = [("threads_per_block", C.c_uint32), ("tuples_per_thread", C.c_uint32),
                ("num_blocks", C.c_uint32), ("flags", C.c_uint32)]


class Timing(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("h2d_ms", C.c_double), ("wall_ms", C.c_double),
                ("bytes_read", C.c_uint64), ("rows", C.c_uint64), ("num_blocks", C.c_uint32),
                ("threads_per_block", C.c_uint32), ("tuples_per_thread", C.c_uint32),
                ("placement", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("clock_khz", C.c_int32), ("memory_clock_khz", C.c_int32),
                ("memory_bus_width", C.c_int32), ("total_memory", C.c_uint64),
                ("lds_per_block", C.c_uint64), ("wavefront_size", C.c_int32),
                ("max_threads_per_block", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("table", C.c_void_p), ("row_begin", C.c_uint64), ("row_count", C.c_uint64),
                ("cutoff_day", C.c_int32), ("reserved", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("max_inflight", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("batch_window_us", C.c_uint32),
                ("memory_budget", C.c_uint64), ("exclusive_bytes", C.c_uint64)]


class RequestStats(C.Structure):
    _fields_ = [("submit_to_start_us", C.c_double), ("latency_us", C.c_double),
                ("batch_size", C.c_uint32), ("batch_id", C.c_uint32)]


def _share_torch_hip_runtime():
    """Makes this process use one HIP runtime.  torch bundles its own libamdhip64 (same
    soname as the system one); if the system copy were loaded first, torch would load a
    second runtime and find no GPU.  Loading torch's copy first lets libq1hip.so bind to it.
    Q1_HIP_RUNTIME=system keeps the system runtime (for processes that never use torch)."""
    if os.environ.get("Q1_HIP_RUNTIME", "") == "system":
        return
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    _share_torch_hip_runtime()
    lib = C.CDLL(LIB_PATH)
    P, U64, I32, I = C.c_void_p, C.c_uint64, C.c_int32, C.c_int
    sig = {
        "q1_strerror": (C.c_char_p, [I]),
        "q1_abi_version": (I, []),
        "q1_last_error": (C.c_char_p, [P]),
        "q1_init": (I, [I, C.POINTER(P)]),
        "q1_shutdown": (None, [P]),
        "q1_device_info_get": (I, [P, C.POINTER(DeviceInfo)]),
        "q1_synchronize": (I, [P]),
        "q1_stream": (P, [P]),
        "q1_table_upload": (I, [P, I, C.POINTER(Columns), U64, C.POINTER(P)]),
        "q1_table_wrap": (I, [P, I, C.POINTER(Columns), U64, C.POINTER(P)]),
        "q1_table_generate": (I, [P, I, U64, U64, C.POINTER(P)]),
        "q1_table_info": (I, [P, C.POINTER(I), C.POINTER(U64), C.POINTER(Columns)]),
        "q1_table_download": (I, [P, P, C.POINTER(Columns)]),
        "q1_table_device_bytes": (U64, [P]),
        "q1_table_free": (None, [P]),
        "q1_build_filter": (I, [P, P, I32, C.POINTER(P)]),
        "q1_run": (I, [P, P, U64, U64, I32, I, C.POINTER(Launch), C.POINTER(Result)]),
        "q1_last_timing": (I, [P, C.POINTER(Timing)]),
        "q1_enqueue": (I, [P, I, I, C.POINTER(Launch), P, C.POINTER(Columns), U64, U64, I32, P]),
        "q1_enqueue_noalloc": (I, [P, I, I, C.POINTER(Launch), P, C.POINTER(Columns), U64, U64, I32, P]),
        "q1_table_refresh_domain": (I, [P, P]),
        "q1_host_columns_in_tpch_domain": (I, [I, C.POINTER(Columns), U64, I, C.POINTER(C.c_int)]),
        "q1_aggregates_to_result": (None, [C.POINTER(DeviceAggregates), C.POINTER(Result)]),
        "q1_result_merge": (None, [C.POINTER(Result), C.POINTER(Result), C.POINTER(Result)]),
        "q1_host_alloc": (P, [P, C.c_size_t]),
        "q1_host_free": (None, [P, P]),
        "q1_layout_column_bytes": (I, [I, U64, C.POINTER(U64)]),
        "q1_generate_host": (I, [I, U64, U64, U64, C.POINTER(Columns), I]),
        "q1_compress_columns": (I, [C.POINTER(Columns), U64, C.POINTER(Columns), I]),
        "q1_cpu_build_filter": (I, [P, U64, I32, P, I]),
        "q1_day_number": (I32, [I32, I32, I32]),
        "q1_cpu_run": (I, [I, C.POINTER(Columns), U64, U64, I32, I, C.POINTER(Result)]),
        "q1_run_split": (I, [P, P, C.POINTER(Columns), U64, I32, I, C.POINTER(Launch), C.c_double,
                             U64, I, C.POINTER(Result), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "q1_run_streaming": (I, [P, I, C.POINTER(Columns), U64, I32, I, C.POINTER(Launch), U64, I,
                                 C.POINTER(Result)]),
        "q1_tuning_set": (I, [P, I, I, C.POINTER(Launch)]),
        "q1_tuning_get": (I, [P, I, I, C.POINTER(Launch)]),
        "q1_tuning_resolve": (I, [P, I, I, U64, C.POINTER(Launch)]),
        "q1_tuning_set_best_placement": (I, [P, I, I]),
        "q1_tuning_get_streaming": (I, [P, I, C.POINTER(U64), C.POINTER(I)]),
    }
    optional = {
        "q1_run_batch": (I, [P, C.POINTER(Request), C.c_uint32, C.POINTER(Launch), C.POINTER(Result)]),
        "q1_server_create": (I, [P, C.POINTER(ServerConfig), C.POINTER(P)]),
        "q1_server_submit": (I, [P, C.POINTER(Request), C.POINTER(U64)]),
        "q1_server_wait": (I, [P, U64, C.c_int64, C.POINTER(Result), C.POINTER(RequestStats)]),
        "q1_server_cancel": (I, [P, U64]),
        "q1_server_drain": (I, [P]),
        "q1_server_destroy": (None, [P]),
    }
    for name, (res, args) in list(sig.items()) + list(optional.items()):
        fn = getattr(lib, name, None)
        if fn is None:
            if name in sig:
                raise ImportError(f"{LIB_PATH} does not export {name}")
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def day_number(year: int, month: int, day: int) -> int:
    """MonetDB day number of a date (the numbering of cpu/monetdb.cpp todate)."""
    return lib.q1_day_number(year, month, day)


DEFAULT_CUTOFF = 729999  # 1998-12-01 minus 90 days (cpu/common.cpp:59-60)


def column_bytes(layout: int, rows: int) -> list[int]:
    out = (C.c_uint64 * 8)()
    _check(lib.q1_layout_column_bytes(layout, rows, out), "q1_layout_column_bytes")
    return list(out)


def _check(status: int, what: str, ctx=None):
    if status != OK:
        detail = ""
        if ctx is not None:
            detail = lib.q1_last_error(ctx).decode(errors="replace")
        raise Q1Error(status, f"{what}: {lib.q1_strerror(status).decode()}" + (f" ({detail})" if detail else ""))


def empty_columns(layout: int, rows: int) -> dict:
    """Host numpy arrays of the layout's column shapes (packed columns as uint32 words)."""
    sizes = column_bytes(layout, rows)
    cols = {}
    for i, name in enumerate(COLUMN_NAMES):
        dt = _DTYPES[layout][i]
        if dt is None:
            continue
        n = sizes[i] // np.dtype(dt).itemsize
        cols[name] = np.zeros(n, dtype=dt)
    return cols


def columns_struct(cols: dict) -> Columns:
    c = Columns()
    for name in COLUMN_NAMES:
        arr = cols.get(name)
        if arr is not None:
            assert arr.flags["C_CONTIGUOUS"]
            setattr(c, name, arr.ctypes.data)
    return c


def generate_host(layout: int, seed: int, rows: int, row_begin: int = 0, threads: int = 0) -> dict:
    """Host columns of rows [row_begin, row_begin+rows) of the seeded generator."""
    return fill_host(layout, seed, empty_columns(layout, rows), rows, row_begin, threads)


def fill_host(layout: int, seed: int, cols: dict, rows: int, row_begin: int = 0, threads: int = 0) -> dict:
    """Writes rows [row_begin, row_begin+rows) of the seeded generator into existing host
    columns of the layout's shapes (e.g. pinned ones from Context.pinned_columns)."""
    st = columns_struct(cols)
    _check(lib.q1_generate_host(layout, seed, row_begin, rows, C.byref(st), threads), "q1_generate_host")
    return cols


def compress_columns(ref_cols: dict, rows: int, threads: int = 0) -> dict:
    out = empty_columns(COMPRESSED, rows)
    a, b = columns_struct(ref_cols), columns_struct(out)
    _check(lib.q1_compress_columns(C.byref(a), rows, C.byref(b), threads), "q1_compress_columns")
    return out


def cpu_build_filter(ship_u16: np.ndarray, cutoff_day: int, threads: int = 0) -> np.ndarray:
    rows = len(ship_u16)
    out = np.zeros((rows + 31) // 32, dtype=np.uint32)
    _check(lib.q1_cpu_build_filter(ship_u16.ctypes.data, rows, cutoff_day, out.ctypes.data, threads),
           "q1_cpu_build_filter")
    return out


def pushdown_columns(compressed_cols: dict, cutoff_day: int, rows: int) -> dict:
    cols = {k: v for k, v in compressed_cols.items() if k != "ship_date"}
    cols["filter"] = cpu_build_filter(compressed_cols["ship_date"][:rows], cutoff_day)
    return cols


@dataclass
class GroupResult:
    sum_qty: int
    sum_base_price: int
    sum_disc_price: int
    sum_charge: int
    sum_disc: int
    count: int


def _i128(lo: int, hi: int) -> int:
    return (int(hi) << 64) + int(lo)


def result_to_list(r: Result) -> list[GroupResult]:
    out = []
    for g in r.g:
        out.append(GroupResult(int(g.sum_qty), int(g.sum_base_price),
                               _i128(g.sum_disc_price_lo, g.sum_disc_price_hi),
                               _i128(g.sum_charge_lo, g.sum_charge_hi), int(g.sum_disc), int(g.count)))
    return out


def cpu_run(layout: int, cols: dict, row_begin: int, row_count: int, cutoff_day: int = DEFAULT_CUTOFF,
            threads: int = 0) -> list[GroupResult]:
    """The CPU coprocessor's aggregation of host columns (morsel-parallel)."""
    st = columns_struct(cols)
    r = Result()
    _check(lib.q1_cpu_run(layout, C.byref(st), row_begin, row_count, cutoff_day, threads, C.byref(r)),
           "q1_cpu_run")
    return result_to_list(r)


def make_launch(threads_per_block=0, tuples_per_thread=0, num_blocks=0, flags=0):
    return Launch(threads_per_block, tuples_per_thread, num_blocks, flags)


class Table:
    """A device-resident table (owned by the library)."""

    def __init__(self, ctx: "Context", handle: int, layout: int, rows: int, keep=None):
        self.ctx, self.handle, self.layout, self.rows = ctx, handle, layout, rows
        self._keep = keep  # parent table of a pushdown table

    def free(self):
        if self.handle:
            lib.q1_table_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def device_bytes(self) -> int:
        return int(lib.q1_table_device_bytes(self.handle))

    def download(self) -> dict:
        cols = empty_columns(self.layout, self.rows)
        st = columns_struct(cols)
        self.ctx.check(lib.q1_table_download(self.ctx.handle, self.handle, C.byref(st)), "q1_table_download")
        return cols

    def device_columns(self) -> Columns:
        c = Columns()
        self.ctx.check(lib.q1_table_info(self.handle, None, None, C.byref(c)), "q1_table_info")
        return c


class Context:
    """One device context (stream, events, scratch); single-threaded like the reference's
    execute_query_1_once caller."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib.q1_init(device, C.byref(h))
        if st != OK:
            raise Q1Error(st, "q1_init: " + lib.q1_strerror(st).decode())
        self.handle = h.value

    def check(self, status: int, what: str):
        _check(status, what, self.handle)

    def close(self):
        if self.handle:
            lib.q1_shutdown(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self) -> DeviceInfo:
        d = DeviceInfo()
        self.check(lib.q1_device_info_get(self.handle, C.byref(d)), "q1_device_info_get")
        return d

    def stream(self) -> int:
        return lib.q1_stream(self.handle)

    def synchronize(self):
        self.check(lib.q1_synchronize(self.handle), "q1_synchronize")

    # --- tables -----------------------------------------------------------------------
    def upload(self, layout: int, cols: dict, rows: int) -> Table:
        st = columns_struct(cols)
        h = C.c_void_p()
        self.check(lib.q1_table_upload(self.handle, layout, C.byref(st), rows, C.byref(h)), "q1_table_upload")
        return Table(self, h.value, layout, rows)

    def wrap(self, layout: int, dev_cols: Columns, rows: int) -> Table:
        h = C.c_void_p()
        self.check(lib.q1_table_wrap(self.handle, layout, C.byref(dev_cols), rows, C.byref(h)), "q1_table_wrap")
        return Table(self, h.value, layout, rows)

    def generate(self, layout: int, seed: int, rows: int) -> Table:
        h = C.c_void_p()
        self.check(lib.q1_table_generate(self.handle, layout, seed, rows, C.byref(h)), "q1_table_generate")
        return Table(self, h.value, layout, rows)

    def build_filter(self, table: Table, cutoff_day: int) -> Table:
        h = C.c_void_p()
        self.check(lib.q1_build_filter(self.handle, table.handle, cutoff_day, C.byref(h)), "q1_build_filter")
        return Table(self, h.value, COMPRESSED_PUSHDOWN, table.rows, keep=table)

    # --- queries ----------------------------------------------------------------------
    def run(self, table: Table, row_begin: int = 0, row_count: int | None = None,
            cutoff_day: int = DEFAULT_CUTOFF, placement=IN_REGISTERS, launch: Launch | None = None):
        if isinstance(placement, str):
            placement = PLACEMENTS[placement]
        if row_count is None:
            row_count = table.rows - row_begin
        r = Result()
        self.check(lib.q1_run(self.handle, table.handle, row_begin, row_count, cutoff_day, placement,
                              C.byref(launch) if launch is not None else None, C.byref(r)), "q1_run")
        return result_to_list(r)

    def run_raw(self, table: Table, row_begin, row_count, cutoff_day, placement, launch=None) -> Result:
        r = Result()
        self.check(lib.q1_run(self.handle, table.handle, row_begin, row_count, cutoff_day, placement,
                              C.byref(launch) if launch is not None else None, C.byref(r)), "q1_run")
        return r

    def last_timing(self) -> Timing:
        t = Timing()
        self.check(lib.q1_last_timing(self.handle, C.byref(t)), "q1_last_timing")
        return t

    def run_split(self, table: Table, host_cols: dict, rows: int, cpu_fraction: float,
                  cutoff_day: int = DEFAULT_CUTOFF, placement=IN_REGISTERS, launch=None,
                  align_rows: int = 0, cpu_threads: int = 0):
        st = columns_struct(host_cols)
        r = Result()
        cms, gms = C.c_double(), C.c_double()
        self.check(lib.q1_run_split(self.handle, table.handle, C.byref(st), rows, cutoff_day, placement,
                                    C.byref(launch) if launch is not None else None, cpu_fraction,
                                    align_rows, cpu_threads, C.byref(r), C.byref(cms), C.byref(gms)),
                   "q1_run_split")
        return result_to_list(r), cms.value, gms.value

    def run_streaming(self, layout: int, host_cols: dict, rows: int, cutoff_day: int = DEFAULT_CUTOFF,
                      placement=IN_REGISTERS, launch=None, chunk_rows: int = 0, streams: int = 0):
        st = columns_struct(host_cols)
        r = Result()
        self.check(lib.q1_run_streaming(self.handle, layout, C.byref(st), rows, cutoff_day, placement,
                                        C.byref(launch) if launch is not None else None, chunk_rows,
                                        streams, C.byref(r)), "q1_run_streaming")
        return result_to_list(r)

    def tuning_set(self, layout: int, placement: int, tpb: int, tpt: int):
        l = Launch(tpb, tpt, 0, 0)
        self.check(lib.q1_tuning_set(self.handle, layout, placement, C.byref(l)), "q1_tuning_set")

    def tuning_get(self, layout: int, placement: int) -> Launch:
        l = Launch()
        self.check(lib.q1_tuning_get(self.handle, layout, placement, C.byref(l)), "q1_tuning_get")
        return l

    def tuning_resolve(self, layout: int, placement: int, rows: int) -> Launch:
        """The shape a q1_run of `rows` rows with no caller shape takes."""
        l = Launch()
        self.check(lib.q1_tuning_resolve(self.handle, layout, placement, rows, C.byref(l)), "q1_tuning_resolve")
        return l

    def tuning_get_streaming(self, layout: int):
        """(chunk_rows, streams) that q1_run_streaming takes when given 0."""
        chunk, streams = C.c_uint64(), C.c_int()
        self.check(lib.q1_tuning_get_streaming(self.handle, layout, C.byref(chunk), C.byref(streams)),
                   "q1_tuning_get_streaming")
        return chunk.value, streams.value

    def set_best_placement(self, layout: int, placement: int):
        self.check(lib.q1_tuning_set_best_placement(self.handle, layout, placement),
                   "q1_tuning_set_best_placement")

    def host_alloc(self, nbytes: int) -> int:
        p = lib.q1_host_alloc(self.handle, nbytes)
        if not p:
            raise Q1Error(ENOMEM, "q1_host_alloc")
        return p

    def host_free(self, p: int):
        lib.q1_host_free(self.handle, p)

    def pinned_columns(self, layout: int, rows: int) -> dict:
        """Host numpy columns of the layout's shapes in pinned memory (q1_host_alloc), for
        q1_run_streaming and q1_run_split; release them with free_pinned_columns."""
        sizes = column_bytes(layout, rows)
        cols = {}
        try:
            for i, name in enumerate(COLUMN_NAMES):
                dt = _DTYPES[layout][i]
                if dt is None or sizes[i] == 0:
                    continue
                buf = (C.c_uint8 * sizes[i]).from_address(self.host_alloc(sizes[i]))
                cols[name] = np.frombuffer(buf, dtype=dt)
        except Exception:
            self.free_pinned_columns(cols)
            raise
        return cols

    def free_pinned_columns(self, cols: dict):
        for arr in cols.values():
            self.host_free(arr.ctypes.data)
        cols.clear()
