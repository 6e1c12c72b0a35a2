"""Persistent tuning cache over `Physics.autotune`.

The choices the library tunes (candidate-scan margin, LDS staging pool, row
sort width, dense-scan variant) never
change results, only speed, and the best one depends on the scene: its size,
grid, spring mode and how unevenly the spheres fill the cells.  This module
keys a JSON cache on those, so a scene type is timed once per machine.
"""
from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

from ._lib import MbError
from .physics import Physics, PhysicsSettings

DEFAULT_CACHE = Path(os.environ.get("MB_TUNING_CACHE",
                                    Path.home() / ".cache" / "bevy_millions_ball_amd" / "tuning.json"))


def density_class(settings: PhysicsSettings, position_with_scale: np.ndarray) -> str:
    """"uniform" when the fullest cell holds at most 8x the mean occupancy
    (and at most 16 spheres), else "dense"."""
    g = int(settings.number_of_grids_one_dimension)
    gs = float(settings.grid_size)
    hms = float(settings.half_map_size())
    pos = np.asarray(position_with_scale, np.float32).reshape(-1, 4)
    cx = np.clip(((pos[:, 0] + hms) / gs).astype(np.int64), 0, g - 1)
    cz = np.clip(((pos[:, 2] + hms) / gs).astype(np.int64), 0, g - 1)
    counts = np.bincount(cz * g + cx, minlength=g * g)
    mean = max(pos.shape[0] / (g * g), 1.0)
    return "uniform" if counts.max() <= max(8.0 * mean, 16.0) else "dense"


def scene_key(settings: PhysicsSettings, position_with_scale: np.ndarray) -> str:
    return "n{}-g{}-gs{}-{}-{}".format(settings.max_number_of_agents(),
                                       settings.number_of_grids_one_dimension,
                                       settings.grid_size,
                                       "spring" if settings.enable_spring else "rigid",
                                       density_class(settings, position_with_scale))


class TuningCache:
    def __init__(self, path: Optional[Path] = None):
        self.path = Path(path) if path is not None else DEFAULT_CACHE
        self.entries = {}
        if self.path.exists():
            try:
                self.entries = json.loads(self.path.read_text())
            except (OSError, ValueError):
                self.entries = {}

    def get(self, key: str) -> Optional[Tuple[float, int, int, int, int]]:
        e = self.entries.get(key)
        if not e:
            return None
        return (float(e["scan_margin"]), int(e["stage_pool"]),
                int(e.get("row_sort_threads", 512)), int(e.get("dense_scan", 0)),
                int(e.get("one_launch", 0)))

    def put(self, key: str, tuning: Tuple[float, int, int, int, int], ms_per_frame: float) -> None:
        self.entries[key] = {"scan_margin": tuning[0], "stage_pool": tuning[1],
                             "row_sort_threads": tuning[2], "dense_scan": tuning[3],
                             "one_launch": tuning[4], "ms_per_frame": ms_per_frame}
        self.path.parent.mkdir(parents=True, exist_ok=True)
        tmp = self.path.with_suffix(".tmp")
        tmp.write_text(json.dumps(self.entries, indent=1, sort_keys=True))
        os.replace(tmp, self.path)


def tune(physics: Physics, cache: Optional[TuningCache] = None,
         frames_per_trial: int = 5) -> Tuple[Tuple[float, int, int, int, int], bool]:
    """Applies the cached tuning for the current scene, or times the candidates
    and caches the fastest.  Returns (tuning, came_from_cache)."""
    cache = cache if cache is not None else TuningCache()
    pos, _ = physics.download_state()
    key = scene_key(physics.settings, pos)
    hit = cache.get(key)
    if hit is not None:
        try:
            physics.set_tuning(*hit)
            return hit, True
        except MbError:
            pass  # an entry from a library with other tuning choices: tune again
    best, ms = physics.autotune(frames_per_trial)
    cache.put(key, best, ms)
    return best, False
