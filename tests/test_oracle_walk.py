"""The oracle's neighbour walk (physics.wgsl:119-206) on grids of several
cells: closed-form pairs across each neighbour direction and the grid's edges,
non-interaction two cells apart, and the visited candidate list of every slot
against a brute-force restatement of the 3 x 3 neighbourhood (tests/
walk_checks.py says how each is built).  test_oracle_mutants.py shows that
copies of the oracle with a piece of the walk removed fail these checks."""
import pytest

import walk_checks as W


@pytest.mark.parametrize("check", W.CHECKS, ids=lambda c: c.__name__)
def test_walk(oracle, check):
    check(oracle)
