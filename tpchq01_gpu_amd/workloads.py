"""The synthetic workloads the path is measured and checked on, in one place so that the
benchmark and the tests build exactly the same tables and requests.

  * scan:    600,000,000 device-resident rows, seed 0x51C00002 (compressed layout);
  * sample:  6,000,000 and 60,000,000 rows, seed 0x51C00000 (the CPU reference case and
             the uncompressed every-placement case);
  * batch:   256 requests, request r on its own table of seed 0x51C00003 ^ r, row counts
             log-uniform in [1e4, 2e7], cutoff dates uniform in 1998-06-01..1998-11-30;
  * serving: 8 resident compressed tables of 1M..600M rows (geometric), seeds
             0x51C00004 + k, Poisson arrivals with a random table, row range and cutoff.

Every table is the library's seeded generator, so the device copy (q1_table_generate) and
any host copy of any row range (q1_generate_host with row_begin) hold the same rows.
"""
from __future__ import annotations

import numpy as np

SCAN_SEED = 0x51C00002
SCAN_ROWS = 600_000_000
SAMPLE_SEED = 0x51C00000
BATCH_SEED = 0x51C00003
SERVING_SEED = 0x51C00004
CUTOFF_FIRST, CUTOFF_LAST = 729906, 730088     # 1998-06-01 .. 1998-11-30


def batch_spec(n: int = 256):
    """(seed, rows, cutoff) of each request of the batched workload."""
    rng = np.random.default_rng(BATCH_SEED)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(2e7), n)).astype(np.int64)
    cutoffs = rng.integers(CUTOFF_FIRST, CUTOFF_LAST + 1, n)
    return [(BATCH_SEED ^ r, int(sizes[r]), int(cutoffs[r])) for r in range(n)]


def serving_spec():
    """(seed, rows) of the serving workload's 8 resident tables."""
    sizes = np.geomspace(1e6, 6e8, 8).astype(np.int64)
    return [(SERVING_SEED + k, int(sizes[k])) for k in range(8)]
