"""MI355X (gfx950) TPC-H Q1 scan -> filter -> group-aggregate path.

The compute path is the HIP library ``lib/libq1hip.so`` (kernels and C ABI, see
include/q1_hip.h); ``q1`` binds it with ctypes.  ``autotune`` sweeps launch shapes and
``serving`` drives the continuous-batching server.
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC_DIR = os.path.join(PACKAGE_DIR, "csrc")
LIB_DIR = os.path.join(PACKAGE_DIR, "lib")
BIN_DIR = os.path.join(PACKAGE_DIR, "bin")
