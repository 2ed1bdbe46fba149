"""TEST INFRASTRUCTURE ONLY: ctypes loader of the CPU oracle (oracle/q1_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this module,
and only as the test reference / the timed CPU baseline -- never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")


class OGroup(C.Structure):
    _fields_ = [("sum_qty", C.c_int64), ("sum_base_price", C.c_int64),
                ("sum_disc_price_lo", C.c_uint64), ("sum_disc_price_hi", C.c_int64),
                ("sum_charge_lo", C.c_uint64), ("sum_charge_hi", C.c_int64),
                ("sum_disc", C.c_int64), ("count", C.c_int64)]


class OResult(C.Structure):
    _fields_ = [("g", OGroup * 6), ("unexpected_keys", C.c_int64)]


class OLineitem(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ["shipdate", "quantity", "extendedprice", "discount", "tax", "returnflag", "linestatus"]]


class OCompact(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ["shipdate", "returnflag", "linestatus", "discount", "tax", "extendedprice", "quantity"]]


class OCompressed(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ["shipdate", "quantity", "extendedprice", "discount", "tax", "returnflag", "linestatus"]]


COMMON_FLAGS = "-O3 -ftree-vectorize -funroll-all-loops -fomit-frame-pointer"
# ISA-level builds for the timed CPU baseline, best first, with the CPU features each needs.
ISA_BUILDS = [("x86-64-v4", {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}),
              ("x86-64-v3", {"avx2", "bmi2", "fma", "movbe", "f16c"})]


def baseline_build() -> tuple[str, str]:
    """The oracle build for timing on this host: the highest ISA level whose features
    /proc/cpuinfo lists (the reference builds with -march=native), and its compiler flags."""
    flags = set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        pass
    for name, need in ISA_BUILDS:
        path = os.path.join(BUILD, f"liboracle_{name}.so")
        if need <= flags and os.path.exists(path):
            return path, f"gcc {COMMON_FLAGS} -march={name}"
    return os.path.join(BUILD, "liboracle.so"), f"gcc {COMMON_FLAGS}"


def build(mutant: int | None = None) -> str:
    """Builds the oracle library (or a planted-mistake variant) and returns its path."""
    if mutant is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        return os.path.join(BUILD, "liboracle.so")
    subprocess.run(["make", "-s", "-C", HERE, "mutant", f"MUTANT={mutant}"], check=True)
    return os.path.join(BUILD, f"liboracle_mutant{mutant}.so")


class Oracle:
    def __init__(self, path: str | None = None):
        if path is None:
            path = os.path.join(BUILD, "liboracle.so")
            if not os.path.exists(path):
                path = build()
        self.path = path
        lib = C.CDLL(path)
        P, I64, I32 = C.c_void_p, C.c_int64, C.c_int32
        lib.q1o_todate.restype = I32
        lib.q1o_todate.argtypes = [C.c_int, C.c_int, C.c_int]
        for name in ("q1o_naive",):
            getattr(lib, name).restype = None
            getattr(lib, name).argtypes = [C.POINTER(OLineitem), I64, I64, I32, C.POINTER(OResult)]
        lib.q1o_hyper.restype = C.c_int
        lib.q1o_hyper.argtypes = [C.POINTER(OLineitem), I64, I64, I32, C.POINTER(OResult)]
        lib.q1o_hyper_parallel.restype = C.c_int
        lib.q1o_hyper_parallel.argtypes = [C.POINTER(OLineitem), I64, I64, I32, C.c_int, I64, C.POINTER(OResult)]
        lib.q1o_naive_parallel.restype = None
        lib.q1o_naive_parallel.argtypes = [C.POINTER(OLineitem), I64, I64, I32, C.c_int, I64, C.POINTER(OResult)]
        lib.q1o_compress.restype = None
        lib.q1o_compress.argtypes = [C.POINTER(OLineitem), I64, C.POINTER(OCompressed)]
        lib.q1o_get_return_flag.restype = C.c_uint32
        lib.q1o_get_return_flag.argtypes = [P, I64]
        lib.q1o_get_line_status.restype = C.c_uint32
        lib.q1o_get_line_status.argtypes = [P, I64]
        lib.q1o_compressed_kernel.restype = None
        lib.q1o_compressed_kernel.argtypes = [C.POINTER(OCompressed), I64, I64, I32, C.POINTER(OResult)]
        lib.q1o_compressed_parallel.restype = None
        lib.q1o_compressed_parallel.argtypes = [C.POINTER(OCompressed), I64, I64, I32, C.c_int, I64,
                                                C.POINTER(OResult)]
        lib.q1o_precompute_filter.restype = None
        lib.q1o_precompute_filter.argtypes = [P, I64, I32, P]
        lib.q1o_pushdown_kernel.restype = None
        lib.q1o_pushdown_kernel.argtypes = [P, C.POINTER(OCompressed), I64, I64, C.POINTER(OResult)]
        lib.q1o_compact_init.restype = None
        lib.q1o_compact_init.argtypes = [C.POINTER(OLineitem), I64, C.POINTER(OCompact)]
        lib.q1o_x100.restype = None
        lib.q1o_x100.argtypes = [C.POINTER(OCompact), I64, I64, I32, C.c_int, I64, C.POINTER(OResult)]
        lib.q1o_avx512.restype = None
        lib.q1o_avx512.argtypes = [C.POINTER(OLineitem), I64, I64, I32, C.POINTER(OResult)]
        self.lib = lib

    # ---- helpers ---------------------------------------------------------------------
    @staticmethod
    def _li(cols: dict) -> OLineitem:
        li = OLineitem()
        li.shipdate = cols["ship_date"].ctypes.data
        li.quantity = cols["quantity"].ctypes.data
        li.extendedprice = cols["extended_price"].ctypes.data
        li.discount = cols["discount"].ctypes.data
        li.tax = cols["tax"].ctypes.data
        li.returnflag = cols["return_flag"].ctypes.data
        li.linestatus = cols["line_status"].ctypes.data
        return li

    @staticmethod
    def _cc(cols: dict) -> OCompressed:
        c = OCompressed()
        c.shipdate = cols["ship_date"].ctypes.data if "ship_date" in cols else None
        c.quantity = cols["quantity"].ctypes.data
        c.extendedprice = cols["extended_price"].ctypes.data
        c.discount = cols["discount"].ctypes.data
        c.tax = cols["tax"].ctypes.data
        c.returnflag = cols["return_flag"].ctypes.data
        c.linestatus = cols["line_status"].ctypes.data
        return c

    @staticmethod
    def _check_ref(cols: dict):
        assert cols["ship_date"].dtype == np.int32
        for k in ("quantity", "extended_price", "discount", "tax"):
            assert cols[k].dtype == np.int64, k
        for k in ("return_flag", "line_status"):
            assert cols[k].dtype.itemsize == 1, k

    @staticmethod
    def to_list(r: OResult) -> list[dict]:
        out = []
        for g in r.g:
            out.append(dict(sum_qty=int(g.sum_qty), sum_base_price=int(g.sum_base_price),
                            sum_disc_price=(int(g.sum_disc_price_hi) << 64) + int(g.sum_disc_price_lo),
                            sum_charge=(int(g.sum_charge_hi) << 64) + int(g.sum_charge_lo),
                            sum_disc=int(g.sum_disc), count=int(g.count)))
        return out

    # ---- the oracle functions ------------------------------------------------------------
    def todate(self, day: int, month: int, year: int) -> int:
        return self.lib.q1o_todate(day, month, year)

    def naive(self, ref_cols: dict, begin: int, end: int, cutoff: int):
        self._check_ref(ref_cols)
        r = OResult()
        self.lib.q1o_naive(C.byref(self._li(ref_cols)), begin, end, cutoff, C.byref(r))
        return self.to_list(r), int(r.unexpected_keys)

    def hyper(self, ref_cols: dict, begin: int, end: int, cutoff: int):
        self._check_ref(ref_cols)
        r = OResult()
        flag = self.lib.q1o_hyper(C.byref(self._li(ref_cols)), begin, end, cutoff, C.byref(r))
        return self.to_list(r), int(flag)

    def hyper_parallel(self, ref_cols: dict, begin: int, end: int, cutoff: int, threads: int,
                       morsel_rows: int = 10 * 1024):
        self._check_ref(ref_cols)
        r = OResult()
        flag = self.lib.q1o_hyper_parallel(C.byref(self._li(ref_cols)), begin, end, cutoff, threads, morsel_rows,
                                           C.byref(r))
        return self.to_list(r), int(flag)

    def naive_parallel(self, ref_cols: dict, begin: int, end: int, cutoff: int, threads: int,
                       morsel_rows: int = 10 * 1024):
        self._check_ref(ref_cols)
        r = OResult()
        self.lib.q1o_naive_parallel(C.byref(self._li(ref_cols)), begin, end, cutoff, threads, morsel_rows,
                                    C.byref(r))
        return self.to_list(r), int(r.unexpected_keys)

    def compress(self, ref_cols: dict, n: int) -> dict:
        self._check_ref(ref_cols)
        out = {
            "ship_date": np.zeros(n, np.uint16), "quantity": np.zeros(n, np.uint8),
            "extended_price": np.zeros(n, np.uint32), "discount": np.zeros(n, np.uint8),
            "tax": np.zeros(n, np.uint8),
            "return_flag": np.zeros(max((n + 15) // 16, 1), np.uint32),
            "line_status": np.zeros(max((n + 31) // 32, 1), np.uint32),
        }
        self.lib.q1o_compress(C.byref(self._li(ref_cols)), n, C.byref(self._cc(out)))
        out["return_flag"] = out["return_flag"][: (n + 15) // 16]
        out["line_status"] = out["line_status"][: (n + 31) // 32]
        return out

    def get_return_flag(self, words: np.ndarray, row: int) -> int:
        return int(self.lib.q1o_get_return_flag(words.ctypes.data, row))

    def get_line_status(self, words: np.ndarray, row: int) -> int:
        return int(self.lib.q1o_get_line_status(words.ctypes.data, row))

    def compressed_kernel(self, ccols: dict, begin: int, end: int, cutoff: int):
        r = OResult()
        self.lib.q1o_compressed_kernel(C.byref(self._cc(ccols)), begin, end, cutoff, C.byref(r))
        return self.to_list(r)

    def compressed_parallel(self, ccols: dict, begin: int, end: int, cutoff: int, threads: int,
                            morsel_rows: int = 1 << 20):
        r = OResult()
        self.lib.q1o_compressed_parallel(C.byref(self._cc(ccols)), begin, end, cutoff, threads, morsel_rows,
                                         C.byref(r))
        return self.to_list(r)

    def precompute_filter(self, ship_u16: np.ndarray, n: int, cutoff: int) -> np.ndarray:
        out = np.zeros(max((n + 31) // 32, 1), np.uint32)
        self.lib.q1o_precompute_filter(ship_u16.ctypes.data, n, cutoff, out.ctypes.data)
        return out[: (n + 31) // 32]

    def pushdown_kernel(self, filt: np.ndarray, ccols: dict, begin: int, end: int):
        r = OResult()
        self.lib.q1o_pushdown_kernel(filt.ctypes.data, C.byref(self._cc(ccols)), begin, end, C.byref(r))
        return self.to_list(r)


    def compact(self, ref_cols: dict, n: int) -> dict:
        """ComprData: the compact columns the X100 kernel reads."""
        self._check_ref(ref_cols)
        out = {"ship_date": np.zeros(n, np.int16), "return_flag": np.zeros(n, np.int8),
               "line_status": np.zeros(n, np.int8), "discount": np.zeros(n, np.int8),
               "tax": np.zeros(n, np.int8), "extended_price": np.zeros(n, np.int32),
               "quantity": np.zeros(n, np.int16)}
        c = OCompact(*[out[k].ctypes.data for k in ("ship_date", "return_flag", "line_status", "discount",
                                                     "tax", "extended_price", "quantity")])
        self.lib.q1o_compact_init(C.byref(self._li(ref_cols)), n, C.byref(c))
        return out

    def x100(self, compact_cols: dict, begin: int, end: int, cutoff: int, threads: int = 1,
             morsel_rows: int = 10 * 1024):
        c = OCompact(*[compact_cols[k].ctypes.data for k in ("ship_date", "return_flag", "line_status",
                                                              "discount", "tax", "extended_price", "quantity")])
        r = OResult()
        self.lib.q1o_x100(C.byref(c), begin, end, cutoff, threads, morsel_rows, C.byref(r))
        return self.to_list(r), int(r.unexpected_keys)

    def avx512(self, ref_cols: dict, begin: int, end: int, cutoff: int):
        self._check_ref(ref_cols)
        r = OResult()
        self.lib.q1o_avx512(C.byref(self._li(ref_cols)), begin, end, cutoff, C.byref(r))
        return self.to_list(r), int(r.unexpected_keys)


def ref_to_compressed_scale(groups: list[dict]) -> list[dict]:
    """The compressed kernels sum quantity in whole units (src/main.cu:783); this scales an
    oracle result over compressed columns to the x100 scale of the reference layout."""
    out = []
    for g in groups:
        h = dict(g)
        h["sum_qty"] = g["sum_qty"] * 100
        out.append(h)
    return out
