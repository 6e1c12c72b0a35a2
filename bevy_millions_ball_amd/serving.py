"""Serving many independent scenes on one GPU (the ``mb_server_*`` C ABI).

A :class:`Server` holds scenes of uneven sphere counts resident in device
memory, admits queued requests first come, first served while they fit under a
byte cap, and steps every resident scene in one batched frame per
:meth:`Server.tick`.  Requests that arrive while others run join the next
frame; finished scenes are copied back to host memory on a second stream and
their memory is reused.  Each scene's results are the same alone and in any
mix.  The reference steps one scene per app; this surface has no counterpart
there.

The request stream of the serving workload (:func:`poisson_requests`): Poisson
arrivals, N log-uniform in [4096, 1,048,576], G = next power of two of
ceil(sqrt(N)) cells per dimension, 1-60 LOOP steps, and the plain-step
distributions (radius 0.5, cell size 1, half map height 8, dt 1/60, e = 0.8,
spring off).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from ._lib import (
    MB_ADMIT_FIFO,
    MB_ADMIT_MOST_STEPS_FIRST,
    MB_REQUEST_DONE,
    MB_REQUEST_QUEUED,
    MB_REQUEST_RETIRING,
    MB_REQUEST_RUNNING,
    MbRequestStatus,
    MbServerConfig,
    MbServerStats,
    check,
    lib,
)
from .physics import PhysicsSettings, _ptr
from .scenes import SceneConfig, generate

STATE_NAMES = {
    MB_REQUEST_QUEUED: "queued",
    MB_REQUEST_RUNNING: "running",
    MB_REQUEST_RETIRING: "retiring",
    MB_REQUEST_DONE: "done",
}


@dataclass
class Collected:
    position_with_scale: np.ndarray
    packed_vfme: np.ndarray
    perm: Optional[np.ndarray] = None
    next_keys: Optional[np.ndarray] = None


ADMISSION = {"fifo": MB_ADMIT_FIFO, "most_steps_first": MB_ADMIT_MOST_STEPS_FIRST}


class Server:
    """mb_server_*: independent scenes stepped together in batched frames.

    admission: "fifo" (first come, first served) or "most_steps_first" (the
    queued request with the most LOOP frames goes next)."""

    def __init__(self, memory_cap_bytes: int = 200 * 10**9, chunk_bytes: int = 0,
                 use_graph: bool = False, profile_frames: bool = False,
                 admission: str = "fifo"):
        if admission not in ADMISSION:
            raise ValueError(f"admission must be one of {sorted(ADMISSION)}")
        cfg = MbServerConfig()
        cfg.memory_cap_bytes = int(memory_cap_bytes)
        cfg.chunk_bytes = int(chunk_bytes)
        cfg.use_graph = int(bool(use_graph))
        cfg.profile_frames = int(bool(profile_frames))
        cfg.admission = ADMISSION[admission]
        self._h = C.c_void_p()
        check(lib().mb_server_create(C.byref(cfg), C.byref(self._h)))
        self._n = {}
        self._debug = {}

    @staticmethod
    def scene_bytes(number_of_agents: int, grids: int) -> int:
        return int(lib().mb_server_scene_bytes(int(number_of_agents), int(grids)))

    def close(self) -> None:
        if self._h:
            lib().mb_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def submit(self, settings: PhysicsSettings, position_with_scale: np.ndarray,
               packed_vfme: np.ndarray, steps: int, keep_debug: bool = False) -> int:
        pos = np.ascontiguousarray(position_with_scale, dtype=np.float32).reshape(-1, 4)
        vf = np.ascontiguousarray(packed_vfme, dtype=np.uint32).reshape(-1, 4)
        n = settings.max_number_of_agents()
        if pos.shape[0] != n or vf.shape[0] != n:
            raise ValueError(f"state must hold {n} spheres")
        cs = settings.to_c()
        rid = C.c_uint64()
        check(lib().mb_server_submit(self._h, C.byref(cs), _ptr(pos), _ptr(vf), int(steps),
                                     int(bool(keep_debug)), C.byref(rid)))
        self._n[rid.value] = n
        self._debug[rid.value] = bool(keep_debug)
        return rid.value

    def tick(self) -> int:
        running = C.c_uint32()
        check(lib().mb_server_tick(self._h, C.byref(running)))
        return running.value

    def drain(self) -> None:
        check(lib().mb_server_drain(self._h))

    def status(self, rid: int) -> dict:
        st = MbRequestStatus()
        check(lib().mb_server_status(self._h, C.c_uint64(rid), C.byref(st)))
        return {
            "state": STATE_NAMES[st.state],
            "number_of_agents": st.number_of_agents,
            "steps": st.steps,
            "frames_done": st.frames_done,
            "submit_s": st.submit_s,
            "admit_s": st.admit_s,
            "done_s": st.done_s,
        }

    def collect(self, rid: int) -> Collected:
        n = self._n[rid]
        pos = np.empty((n, 4), np.float32)
        vf = np.empty((n, 4), np.uint32)
        perm = keys = None
        if self._debug[rid]:
            perm = np.empty(n, np.uint32)
            keys = np.empty(n, np.uint32)
        check(lib().mb_server_collect(self._h, C.c_uint64(rid), _ptr(pos), _ptr(vf),
                                      _ptr(perm) if perm is not None else None,
                                      _ptr(keys) if keys is not None else None))
        del self._n[rid], self._debug[rid]
        return Collected(pos, vf, perm, keys)

    def stats(self) -> dict:
        st = MbServerStats()
        check(lib().mb_server_stats_read(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in MbServerStats._fields_}

    def frame_times(self) -> np.ndarray:
        count = C.c_uint32()
        check(lib().mb_server_frame_times(self._h, None, 0, C.byref(count)))
        out = np.empty(max(count.value, 1), np.float64)
        check(lib().mb_server_frame_times(self._h, _ptr(out), count.value, C.byref(count)))
        return out[: count.value]


def scene_grid(number_of_agents: int) -> int:
    """Next power of two of ceil(sqrt(N))."""
    side = math.isqrt(number_of_agents - 1) + 1 if number_of_agents > 1 else 1
    return 1 << (side - 1).bit_length()


def serving_scene_config(number_of_agents: int) -> SceneConfig:
    g = scene_grid(number_of_agents)
    return SceneConfig(f"serve_{number_of_agents}", number_of_agents, g)


@dataclass
class Request:
    arrival_s: float
    number_of_agents: int
    steps: int
    seed: int


def poisson_requests(count: int, rate_per_s: float = 200.0, seed: int = 0,
                     n_min: int = 4096, n_max: int = 1 << 20,
                     steps_min: int = 1, steps_max: int = 60) -> List[Request]:
    """`count` requests of the serving workload, seeded (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    ns = np.exp(rng.uniform(math.log(n_min), math.log(n_max), count)).astype(np.int64)
    ns = np.clip(ns, n_min, n_max)
    steps = rng.integers(steps_min, steps_max + 1, count)
    seeds = rng.integers(0, 2**31 - 1, count)
    return [Request(float(a), int(n), int(s), int(sd))
            for a, n, s, sd in zip(arrivals, ns, steps, seeds)]


def request_scene(req: Request):
    """(settings, position_with_scale, packed_vfme) of one serving request."""
    cfg = serving_scene_config(req.number_of_agents)
    pos, vf = generate(cfg, seed=req.seed)
    return cfg.settings(), pos, vf
