"""The oracle is pinned before it is trusted (see tests/pinning.py for what pins what)."""
import pytest

import pinning
from oracle_lib import Oracle, build

# Planted mistakes in oracle/q1_oracle.c (-DQ1O_MUTANT=k) and the check meant to catch each.
MUTANTS = {
    1: "brute_force",        # discount and tax swapped
    2: "brute_force",        # tax factor dropped
    3: "brute_force",        # discount added
    4: "filter_boundary",    # strict filter
    5: "brute_force",        # flag/status swapped in the slot
    6: "compress",           # return-flag bit offset
    7: "compress",           # line-status bit index
    8: "int128",             # 64-bit charge sums
    9: "compress",           # quantity not narrowed
    10: "filter_bits",       # last filter word dropped
    11: "brute_force",       # sum_disc of (1 - disc)
    12: "x100",              # X100 drops a morsel's last partial vector
    13: "avx512",            # AVX-512 final sum leaves out one lane
    14: "compressed_parallel",  # compressed morsel-parallel drops a range's last partial morsel
}


@pytest.mark.parametrize("name", sorted(pinning.CHECKS))
def test_oracle_pinned(oracle, name):
    assert pinning.CHECKS[name](oracle), f"oracle fails pinning check {name}"


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_planted_mistake_is_caught(k):
    mutant = Oracle(build(k))
    results = pinning.run_all(mutant)
    failed = [n for n, ok in results.items() if not ok]
    assert failed, f"mutant {k} passed every pinning check"
    assert MUTANTS[k] in failed, f"mutant {k} caught by {failed}, expected {MUTANTS[k]}"
