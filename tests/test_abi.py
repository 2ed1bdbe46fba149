"""The C ABI library loads without a GPU and exports every function include/q1_hip.h
declares; calls that need no device behave (status codes, strings, argument checks)."""
import ctypes as C
import os
import re

from tpchq01_gpu_amd import q1

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "q1_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(q1_[a-z0-9_]+)\s*\(", text))
    # typedef'd struct tags are not functions
    return sorted(n for n in names if not n.endswith("_t"))


def test_every_declared_function_is_exported():
    lib = C.CDLL(q1.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_strings():
    assert q1.lib.q1_abi_version() == 2
    for s in range(8):
        assert q1.lib.q1_strerror(s)
    assert q1.lib.q1_strerror(99) == b"unknown status"


def test_null_arguments_are_rejected():
    assert q1.lib.q1_init(0, None) == q1.EINVAL
    assert q1.lib.q1_run(None, None, 0, 0, 0, 0, None, None) == q1.EINVAL
    assert q1.lib.q1_cpu_run(0, None, 0, 0, 0, 0, None) == q1.EINVAL
    assert q1.lib.q1_layout_column_bytes(9, 10, (C.c_uint64 * 8)()) == q1.EINVAL
    assert q1.lib.q1_generate_host(q1.COMPRESSED, 1, 5, 10, C.byref(q1.Columns()), 1) == q1.EINVAL  # mid-word start


def test_column_bytes():
    assert q1.column_bytes(q1.COMPRESSED, 33) == [66, 33, 132, 33, 33, 12, 8, 0]
    assert q1.column_bytes(q1.COMPRESSED_PUSHDOWN, 33) == [0, 33, 132, 33, 33, 12, 8, 8]
    assert q1.column_bytes(q1.REFERENCE, 2) == [8, 16, 16, 16, 16, 2, 2, 0]
    assert q1.column_bytes(q1.UNCOMPRESSED, 2) == [8, 8, 16, 2, 2, 2, 2, 0]
    # bits per row used for the roofline: 304, 160, 75, 60
    for layout, bits in enumerate(q1.BITS_PER_ROW):
        b = q1.column_bytes(layout, 3200)
        assert sum(b) * 8 == bits * 3200
