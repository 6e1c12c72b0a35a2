"""Checkpoint format, read on the host (mb_checkpoint_inspect needs no device)."""
import struct

import numpy as np
import pytest

from bevy_millions_ball_amd import MbError, PhysicsSettings, inspect_checkpoint
from bevy_millions_ball_amd import _lib


def make_checkpoint(ps: PhysicsSettings, time: int, pos: np.ndarray, vf: np.ndarray) -> bytes:
    """The documented layout, written independently of the library."""
    n = ps.max_number_of_agents()
    head = b"MBCK" + struct.pack("<I", 1) + bytes(ps.to_c()) + struct.pack("<qI", time, n)
    return head + np.ascontiguousarray(pos, np.float32).tobytes() + \
        np.ascontiguousarray(vf, np.uint32).tobytes()


def settings(n=300, g=16):
    return PhysicsSettings(number_of_grids_one_dimension=g, grid_size=2, half_map_height=4,
                           e_agent=0.75, e_envir=0.5, time_delta=1 / 30, _max_number_of_agents=n)


def test_size_and_header_round_trip():
    ps = settings()
    n = ps.max_number_of_agents()
    rng = np.random.default_rng(0)
    data = make_checkpoint(ps, 41, rng.random((n, 4), np.float32),
                           rng.integers(0, 2**31, (n, 4), dtype=np.uint32))
    assert len(data) == _lib.lib().mb_checkpoint_size(n) == 80 + 32 * n
    got, t, count = inspect_checkpoint(data)
    assert (t, count) == (41, n)
    assert bytes(got.to_c()) == bytes(ps.to_c())


@pytest.mark.parametrize("mutate,fragment", [
    (lambda d: b"XBCK" + d[4:], "not a checkpoint"),
    (lambda d: d[:4] + struct.pack("<I", 2) + d[8:], "version"),
    (lambda d: d[:-1], "size"),
    (lambda d: d[:20], "header"),
    (lambda d: d[:10], "header"),
    (lambda d: d[:68] + struct.pack("<q", -2) + d[76:], "time"),
    (lambda d: d[:76] + struct.pack("<I", 299) + d[80:], "disagrees"),
])
def test_rejects_damaged_checkpoints(mutate, fragment):
    ps = settings()
    n = ps.max_number_of_agents()
    data = make_checkpoint(ps, 3, np.zeros((n, 4), np.float32), np.zeros((n, 4), np.uint32))
    with pytest.raises(MbError, match=fragment):
        inspect_checkpoint(mutate(data))
