"""Pins the oracle's f16 conversions (pack2x16float / unpack2x16float,
physics.wgsl:235-253) against two independent implementations: the x86 F16C
instruction (exhaustively, all 2^32 inputs) and numpy's float16 cast."""
import numpy as np
import pytest


def test_unpack_all_halves_match_numpy(oracle):
    h = np.arange(65536, dtype=np.uint32)
    ours = np.array([oracle.f16_to_f32(int(x)) for x in h], np.float32)
    ref = h.astype(np.uint16).view(np.float16).astype(np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(ours), nan)
    assert np.array_equal(ours[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def _sample_floats():
    rng = np.random.default_rng(1)
    halves = np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16)
    finite = halves[np.isfinite(halves)].astype(np.float32)
    srt = np.unique(finite)
    mids = ((srt[:-1].astype(np.float64) + srt[1:].astype(np.float64)) / 2).astype(np.float32)
    mids_u = mids.view(np.uint32)
    near = np.concatenate([mids_u - 1, mids_u, mids_u + 1]).view(np.float32)
    rand = rng.integers(0, 2**32, 200_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    edge = np.array([65504, 65519.99, 65520, 65536, 1e30, 2.0**-24, 2.0**-25, 2.0**-26,
                     3 * 2.0**-26, 2.0**-14, 2.0**-15, -0.0, 0.0], np.float32)
    allv = np.concatenate([finite, near, rand, edge, -edge])
    return allv[np.isfinite(allv)]


def test_pack_matches_numpy_on_ties_and_random(oracle):
    xs = _sample_floats()
    ours = np.fromiter((oracle.f32_to_f16(float(x)) for x in xs), np.uint16, count=xs.size)
    ref = xs.astype(np.float16).view(np.uint16)
    bad = np.nonzero(ours != ref)[0]
    assert bad.size == 0, (xs[bad[:5]], ours[bad[:5]], ref[bad[:5]])


def test_pack_exhaustive_against_f16c(oracle):
    mismatches = oracle.check_f16_against_f16c()
    if mismatches < 0:
        pytest.skip("host CPU lacks F16C")
    assert mismatches == 0
