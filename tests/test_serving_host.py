"""Host-side checks of the serving surface: ABI struct layouts against the C
compiler, the request stream, validation, and the oracle's G^2 > N worlds
(the serving path's scenes may have more cells than spheres)."""
import ctypes as C
import math
import shutil
import subprocess

import numpy as np
import pytest

from bevy_millions_ball_amd import _lib
from bevy_millions_ball_amd.scenes import SceneConfig, generate
from bevy_millions_ball_amd.serving import Server, poisson_requests, scene_grid

from conftest import ROOT

STRUCTS = {
    "mb_settings": _lib.MbSettings,
    "mb_buffers": _lib.MbBuffers,
    "mb_server_config": _lib.MbServerConfig,
    "mb_request_status": _lib.MbRequestStatus,
    "mb_server_stats": _lib.MbServerStats,
    "mb_tuning": _lib.MbTuning,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_layouts_match_the_c_compiler(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "millions_ball.h"',
             "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True,
                               text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_scene_grid_is_next_power_of_two_of_ceil_sqrt():
    for n in [1, 2, 3, 4, 5, 4096, 4097, 5000, 65536, 65537, 1 << 20]:
        g = scene_grid(n)
        side = math.ceil(math.sqrt(n))
        assert g >= side and g & (g - 1) == 0 and (g == 1 or g // 2 < side)
    assert [scene_grid(n) for n in (4096, 5000, 1 << 20)] == [64, 128, 1024]


def test_poisson_requests_are_seeded_and_in_range():
    a = poisson_requests(500, rate_per_s=200.0, seed=3)
    b = poisson_requests(500, rate_per_s=200.0, seed=3)
    assert a == b
    ns = np.array([r.number_of_agents for r in a])
    steps = np.array([r.steps for r in a])
    arr = np.array([r.arrival_s for r in a])
    assert ns.min() >= 4096 and ns.max() <= 1 << 20
    assert steps.min() >= 1 and steps.max() <= 60
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    # mean gap 1/200 s, log-uniform sizes: median near sqrt(4096 * 2^20) = 65536
    assert 0.004 < np.diff(arr).mean() < 0.006
    assert 40000 < np.median(ns) < 110000


def test_scene_bytes_covers_every_array():
    n, g = 100_000, 512
    # pos x2, vfme x2, keys, perm, scratch indices, scratch pairs, spill list
    per_sphere = 16 * 4 + 4 * 3 + 8 + 8
    cap = max(4 * -(-n // g), 512)
    cap = -(-cap // 64) * 64
    # buckets, cell offsets, one packed word per fill / spill counter, and
    # the spread counters of filings in input order (one per 128-byte line)
    want = per_sphere * n + 8 * g * cap + 4 * (g * g + 1) + 4 * (2 * g + 2) + 128 * g
    got = Server.scene_bytes(n, g)
    assert want <= got <= want + 12 * 256


def test_server_rejects_zero_cap():
    with pytest.raises(_lib.MbError, match="memory_cap_bytes"):
        Server(memory_cap_bytes=0)


def brute_ranges(sorted_keys, cells):
    start = np.full(cells, 0xFFFFFFFF, np.uint64)
    close = np.zeros(cells, np.uint64)
    for i, k in enumerate(sorted_keys):
        if start[k] == 0xFFFFFFFF:
            start[k] = i
        close[k] = i
    return start, close


@pytest.mark.parametrize("n,g", [(100, 16), (1, 2), (300, 32)])
def test_oracle_world_with_more_cells_than_spheres(oracle, n, g):
    cfg = SceneConfig(f"o{n}", n, g)
    pos, vf = generate(cfg, seed=n)
    w = oracle.World(cfg.settings())
    w.upload(pos, vf)
    w.step(3)  # INIT + 2 LOOP
    dbg = w.debug()
    keys_in_frame = None
    p2, _ = w.download()
    w.close()
    perm = dbg["perm"]
    assert np.array_equal(np.sort(perm), np.arange(n))
    # the ranges of the last frame cover all G^2 cells
    assert dbg["start"].shape[0] == g * g
    w2 = oracle.World(cfg.settings())
    w2.upload(pos, vf)
    w2.step(2)
    keys_in_frame = w2.debug()["next_keys"]  # the keys the third frame sorted
    w2.close()
    sorted_keys = keys_in_frame[perm]
    assert np.all(np.diff(sorted_keys.astype(np.int64)) >= 0)
    start, close = brute_ranges(sorted_keys, g * g)
    assert np.array_equal(dbg["start"].astype(np.uint64), start)
    valid = start != 0xFFFFFFFF
    assert np.array_equal(dbg["close"].astype(np.uint64)[valid], close[valid])
    hms = (g * cfg.grid_size) // 2
    assert np.all(np.abs(p2[:, [0, 2]]) <= hms + 1e-3)
