"""GPU parity of the one-launch frame (tuning one_launch = 1): every LOOP
frame is a single kernel whose first G blocks sort the rows and whose other
blocks run the collide tiles, each after waiting (bounded) for the rows
around its slots.  Its results must be the oracle's, bit for bit, like the
two-launch frame's; the failure path of its bounded waits is forced from the
host with a wait limit of one poll, so that no wait can hang."""
from dataclasses import replace

import numpy as np
import pytest

from bevy_millions_ball_amd import Physics
from bevy_millions_ball_amd.scenes import CONFIGS, generate

from test_gpu_parity import compare_frames, make_pair

pytestmark = pytest.mark.gpu

ONE = (1.0, 896, 256, 0, 1)          # small pool: rows of up to 2048 spheres in LDS
ONE_DENSE = (0.0625, 1408, 256, 1, 1)  # the spring-clumps point: big pool, dense scan


@pytest.mark.parametrize("name,n,grids,frames,tuning", [
    ("plain_65k", 4096, 64, 40, ONE),
    ("plain_65k", 65_536, 256, 12, ONE),
    ("uniform_1m", 20_000, 128, 20, ONE),       # N not a multiple of the tile
    ("uniform_1m", 10_000, 100, 20, ONE),       # G not a power of two
    ("spring_clumps_1m", 16_384, 128, 40, ONE),
    ("spring_clumps_1m", 16_384, 128, 40, ONE_DENSE),
    ("plain_65k", 256, 4, 30, ONE),             # one tile, four rows
])
def test_one_launch_bit_exact_small(gpu, oracle, name, n, grids, frames, tuning):
    cfg = replace(CONFIGS[name], number_of_agents=n, grids=grids)
    print(f"[one-launch] {name} n={n} G={grids} frames={frames} {tuning}", flush=True)
    g, r = make_pair(oracle, cfg, seed=n + grids)
    with g:
        g.set_tuning(*tuning)
        compare_frames(g, r, frames)
    r.close()


def test_one_launch_one_row(gpu, oracle):
    """G = 1: a single row, sorted by one block, and two tiles waiting for it."""
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=300, grids=1, grid_size=16)
    g, r = make_pair(oracle, cfg, seed=4)
    with g:
        g.set_tuning(*ONE)
        compare_frames(g, r, 30)
    r.close()


@pytest.mark.parametrize("tuning", [ONE, ONE_DENSE])
def test_one_launch_dense_rows_spill_and_fallbacks(gpu, oracle, tuning):
    """Rows far above the mean on one launch: bucket spill (the spill lists of
    the two sets, a tile's spill waiting for row 0's sort), rows sorted in
    global memory and collide blocks that read neighbours from global memory."""
    from scene_helpers import dense_row_scene

    ps, pos, vf = dense_row_scene()
    print(f"[one-launch dense rows] {tuning}", flush=True)
    g = Physics(ps)
    g.upload_state(pos, vf)
    r = oracle.World(ps)
    r.upload(pos, vf)
    with g:
        g.set_tuning(*tuning)
        compare_frames(g, r, 6)
        keys = r.debug()["next_keys"]
        rows = np.bincount(keys // ps.number_of_grids_one_dimension, minlength=64)
        assert rows.max() > 1024  # still spilling at the last frame
    r.close()


@pytest.mark.parametrize("name,frames,tuning", [
    ("uniform_1m", 3, ONE),
    ("spring_clumps_1m", 3, ONE_DENSE),
    ("uniform_4m", 2, ONE),
    ("uniform_4m", 2, (1.0, 1408, 256, 0, 1)),
])
def test_one_launch_bit_exact_full_size(gpu, oracle, name, frames, tuning):
    cfg = CONFIGS[name]
    print(f"[one-launch full] {name} frames={frames} {tuning}", flush=True)
    g, r = make_pair(oracle, cfg, seed=1)
    with g:
        g.set_tuning(*tuning)
        compare_frames(g, r, frames)
    r.close()


@pytest.mark.parametrize("name,frames,tuning", [("uniform_1m", 200, ONE),
                                                ("spring_clumps_1m", 60, ONE_DENSE)])
def test_one_launch_long_run_full_size(gpu, oracle, name, frames, tuning):
    """Hundreds of one-launch frames at full size, compared every 20 frames
    (a chaotic trajectory: one differing bit anywhere would spread)."""
    cfg = CONFIGS[name]
    print(f"[one-launch long] {name} frames={frames}", flush=True)
    g, r = make_pair(oracle, cfg, seed=7)
    with g:
        g.set_tuning(*tuning)
        compare_frames(g, r, frames, check_every=20)
    r.close()


def test_one_launch_alternating_with_two_launches(gpu, oracle):
    """Frames switch between one and two launches (frame 1, which reads the
    INIT filing's counters, on one launch): the bucket sets, spill lists and
    counters either leaves are what the other reads."""
    cfg = replace(CONFIGS["spring_clumps_1m"], number_of_agents=65_536, grids=256, clumps=8)
    g, r = make_pair(oracle, cfg, seed=5)
    with g:
        g.run(1)  # INIT
        r.step(1)
        for k, tuning in enumerate([ONE, (1.0, 896, 512, 0, 0), ONE_DENSE, ONE,
                                    (0.25, 1408, 256, 1, 0), ONE]):
            print(f"[one-launch alt] block {k} {tuning}", flush=True)
            g.set_tuning(*tuning)
            compare_frames(g, r, 3)
    r.close()


def test_one_launch_with_frame_graph_set(gpu, oracle):
    """frame_graph does not capture one-launch frames (each takes a launch
    number of its own as an argument): with both set, frames are launched
    directly and still match the oracle; switching one_launch off mid-run
    hands over to graph replays."""
    cfg = replace(CONFIGS["uniform_1m"], number_of_agents=65_536, grids=256)
    g, r = make_pair(oracle, cfg, seed=3)
    with g:
        g.set_tuning(*ONE)
        g.set_frame_graph(True)
        g.run(2)
        r.step(2)
        for k in range(4):
            if k == 2:
                g.set_tuning(1.0, 896, 512, 0, 0)
                g.set_frame_graph(True)
            g.run(6)
            r.step(6)
            compare_frames(g, r, 1)
    r.close()


def test_one_launch_after_a_grid_change(gpu, oracle):
    """Grid changes between one-launch frames (the launch's block count and
    its rows change with G); the frames after each match the oracle's."""
    cfg = replace(CONFIGS["uniform_1m"], number_of_agents=65_536, grids=256)
    g, r = make_pair(oracle, cfg, seed=8)
    with g:
        g.set_tuning(*ONE)
        compare_frames(g, r, 4)
        for grids, grid_size in ((128, 2), (64, 4), (256, 1)):  # the same 256 x 256 map
            ps = replace(cfg, grids=grids, grid_size=grid_size).settings()
            g.set_settings(ps)
            r.set_settings(ps)
            compare_frames(g, r, 4)
    r.close()


def test_one_launch_wait_limit_reports_and_recovers(gpu, oracle):
    """A wait limit of one poll makes the tiles that find their rows still
    being sorted give up at once (nothing waits): the next synchronising call
    reports MB_ERR_STATE and the handle refuses to step; a fresh upload clears
    it, and with the default limit the frames match the oracle again."""
    cfg = CONFIGS["uniform_1m"]
    pos, vf = generate(cfg, seed=2)
    with Physics(cfg.settings()) as g:
        g.set_tuning(*ONE)
        g.upload_state(pos, vf)
        g.set_wait_limit(1)
        print("[one-launch wait limit] 1 poll", flush=True)
        g.run(4)
        with pytest.raises(Exception, match="one-launch"):
            g.download_state()
        with pytest.raises(Exception, match="no accepted state"):
            g.run(1)
        g.set_wait_limit(0)
        g.upload_state(pos, vf)
        r = oracle.World(cfg.settings())
        r.upload(pos, vf)
        compare_frames(g, r, 3)
        r.close()
