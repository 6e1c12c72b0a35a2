"""Host side of the tuning cache (bevy_millions_ball_amd.autotune)."""
import numpy as np

from bevy_millions_ball_amd.autotune import TuningCache, density_class, scene_key, tune
from bevy_millions_ball_amd.scenes import CONFIGS, generate
from dataclasses import replace

from scene_helpers import dense_row_scene


def test_density_class_tells_uniform_from_clumped():
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=16384, grids=128)
    pos, _ = generate(cfg, seed=0)
    assert density_class(cfg.settings(), pos) == "uniform"
    ps, dense, _ = dense_row_scene()
    assert density_class(ps, dense) == "dense"
    clumps = CONFIGS["spring_clumps_1m"]  # ~40 spheres per cell at the clump centres
    cpos, _ = generate(clumps, seed=0)
    assert density_class(clumps.settings(), cpos) == "dense"


def test_scene_key_names_the_scene():
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=4096, grids=64)
    pos, _ = generate(cfg, seed=0)
    assert scene_key(cfg.settings(), pos) == "n4096-g64-gs1-rigid-uniform"


def test_cache_round_trip(tmp_path):
    path = tmp_path / "t.json"
    c = TuningCache(path)
    assert c.get("k") is None
    c.put("k", (0.5, 1408, 256, 1, 1), 0.123)
    assert TuningCache(path).get("k") == (0.5, 1408, 256, 1, 1)
    # an entry written before the one-launch frame existed reads as one_launch 0
    path.write_text('{"old": {"scan_margin": 1.0, "stage_pool": 896, "ms_per_frame": 0.05}}')
    assert TuningCache(path).get("old") == (1.0, 896, 512, 0, 0)
    path.write_text("not json")
    assert TuningCache(path).get("k") is None  # a damaged cache is ignored


class FakePhysics:
    """Stands in for Physics: the same calls, no device."""

    def __init__(self, settings, pos):
        self.settings = settings
        self._pos = pos
        self.applied = None
        self.autotuned = 0

    def download_state(self):
        return self._pos, None

    def set_tuning(self, margin, pool, sort=512, dense=0, one=0):
        self.applied = (margin, pool, sort, dense, one)

    def autotune(self, frames):
        self.autotuned += 1
        self.applied = (0.25, 1408, 256, 1, 1)
        return self.applied, 1.5


def test_tune_times_once_then_uses_the_cache(tmp_path):
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=4096, grids=64)
    pos, _ = generate(cfg, seed=0)
    cache = TuningCache(tmp_path / "t.json")
    first = FakePhysics(cfg.settings(), pos)
    assert tune(first, cache) == ((0.25, 1408, 256, 1, 1), False) and first.autotuned == 1
    second = FakePhysics(cfg.settings(), pos)
    assert tune(second, TuningCache(tmp_path / "t.json")) == ((0.25, 1408, 256, 1, 1), True)
    assert second.autotuned == 0 and second.applied == (0.25, 1408, 256, 1, 1)
