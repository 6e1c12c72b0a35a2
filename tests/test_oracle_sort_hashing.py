"""Pins the oracle's radix sort and spatial hashing.

* The reference's own known answers (tests/golden/spatial_hashing_kat.json,
  from src/physics/spatial_hashing.rs:551-632).  Parity in the reference only
  picks which EVE/ODD buffer holds the result; the oracle writes to explicit
  outputs, so both parities of a case run the same computation.
* numpy's stable argsort (an independent stable sort) on random keys,
  including pass ranges that leave high bits unsorted.
"""
import json
from pathlib import Path

import numpy as np
import pytest

KAT = json.loads((Path(__file__).parent / "golden" / "spatial_hashing_kat.json").read_text())


@pytest.mark.parametrize("case", KAT["cases"],
                         ids=lambda c: f"n{c['number_of_keys']}-p{c['pass_count']}-{c['read_from_even']}")
def test_reference_known_answers(oracle, case):
    n = case["number_of_keys"]
    keys = (np.arange(n, dtype=np.uint32) % 256).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = oracle.radix_sort(keys, vals, range(0, case["pass_count"]))
    start, close = oracle.spatial_hashing(sk)
    cells = len(case["start"])
    assert start[:cells].tolist() == case["start"]
    assert close[:cells].tolist() == case["close"]
    # the sort itself: stable, so values of each key stay in input order
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(sv, order.astype(np.uint32))


@pytest.mark.parametrize("n,bits,passes", [(1, 8, 1), (5000, 24, 3), (70_001, 20, 3),
                                           (65_536, 32, 4), (40_000, 32, 3), (9_999, 16, 1)])
def test_sort_matches_numpy_stable(oracle, n, bits, passes):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)
    sk, sv = oracle.radix_sort(keys, None, range(0, passes))
    masked = keys & np.uint32((1 << (8 * passes)) - 1 if passes < 4 else 0xFFFFFFFF)
    order = np.argsort(masked, kind="stable").astype(np.uint32)
    assert np.array_equal(sv, order)
    assert np.array_equal(sk, keys[order])


def test_sort_partial_pass_range(oracle):
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 2**24, 30_000, dtype=np.uint64).astype(np.uint32)
    vals = rng.permutation(30_000).astype(np.uint32)
    sk, sv = oracle.radix_sort(keys, vals, range(1, 3))  # bits 8..23 only
    order = np.argsort((keys >> 8) & 0xFFFF, kind="stable")
    assert np.array_equal(sk, keys[order])
    assert np.array_equal(sv, vals[order])


def test_hashing_ranges_match_bruteforce(oracle):
    rng = np.random.default_rng(3)
    n, cells = 20_000, 4096
    keys = np.sort(rng.integers(0, cells, n)).astype(np.uint32)
    start, close = oracle.spatial_hashing(keys, zero_count=n)
    present = np.unique(keys)
    first = np.searchsorted(keys, present, "left")
    last = np.searchsorted(keys, present, "right") - 1
    assert np.array_equal(start[present], first.astype(np.uint32))
    assert np.array_equal(close[present], last.astype(np.uint32))
    empty = np.setdiff1d(np.arange(n), present)
    assert np.all(start[empty] == 0xFFFFFFFF)
