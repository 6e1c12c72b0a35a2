"""Host-side mirror of the reference's physics plugin surface.

``PhysicsPlugin`` / ``PhysicsSettings`` keep the field names, defaults and
accessors of src/physics/mod.rs:53-436; ``Physics`` plays the part of the
render-graph node and bind groups (``PhysicsNode`` mod.rs:746-913,
``PhysicsBindGroup`` mod.rs:614-744).  Every call goes through the C ABI of
libmillions_ball.so (include/millions_ball.h); nothing here computes physics.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, fields
from typing import Optional, Tuple

import numpy as np

from ._lib import MB_INVALID_INDEX, MbBuffers, MbError, MbSettings, MbTuning, check, lib

NUMBER_OF_THREADS_PER_WORKGROUP = 256  # mod.rs:35


@dataclass
class PhysicsPlugin:
    """``PhysicsPlugin`` (mod.rs:53-85) with its defaults (mod.rs:87-109)."""

    enable_simulation: bool = True
    enable_spring: bool = False
    number_of_grids_one_dimension: int = 128
    grid_size: int = 8
    half_map_height: int = 16
    gravity: Tuple[float, float, float] = (0.0, -9.8, 0.0)
    e_envir: float = 0.90
    e_agent: float = 0.95
    k_agent: float = 25.0
    c_agent: float = 0.1
    r_agent: float = 0.9
    max_number_of_agents: int = 16 * 1024
    time_delta: float = 1.0 / 120.0

    def build(self) -> "Physics":
        """``Plugin::build`` (mod.rs:111-125): allocates the device buffers."""
        return Physics(PhysicsSettings.from_plugin(self))


@dataclass
class PhysicsSettings:
    """``PhysicsSettings`` (mod.rs:294-362).

    ``max_number_of_agents`` is private in the reference and read through an
    accessor; it is kept that way.  The reference's ``half_map_height()``
    accessor (f32) shares its name with the u32 field, which Python cannot
    express: it is ``half_map_height_f32()`` here.
    """

    enable_simulation: bool = True
    enable_spring: bool = False
    number_of_grids_one_dimension: int = 128
    grid_size: int = 8
    half_map_height: int = 16
    gravity: Tuple[float, float, float] = (0.0, -9.8, 0.0)
    e_envir: float = 0.90
    e_agent: float = 0.95
    k_agent: float = 25.0
    c_agent: float = 0.1
    r_agent: float = 0.9
    time_delta: float = 1.0 / 120.0
    _max_number_of_agents: int = field(default=16 * 1024, repr=False)

    @classmethod
    def from_plugin(cls, pp: PhysicsPlugin) -> "PhysicsSettings":
        """``impl From<&PhysicsPlugin> for PhysicsSettings`` (mod.rs:402-425)."""
        kw = {f.name: getattr(pp, f.name) for f in fields(pp) if f.name != "max_number_of_agents"}
        return cls(**kw, _max_number_of_agents=pp.max_number_of_agents)

    # accessors, mod.rs:364-400
    def max_number_of_agents(self) -> int:
        return self._max_number_of_agents

    def max_number_of_grids_on_dimension(self) -> int:
        return math.isqrt(self.max_number_of_agents())

    def map_size_u(self) -> int:
        return (self.number_of_grids_one_dimension * self.grid_size) & 0xFFFFFFFF

    def half_map_size_u(self) -> int:
        return self.map_size_u() // 2

    def map_size(self) -> float:
        return float(np.float32(self.map_size_u()))

    def half_map_size(self) -> float:
        return float(np.float32(self.half_map_size_u()))

    def half_map_height_f32(self) -> float:
        return float(np.float32(self.half_map_height))

    def map_height(self) -> float:
        return float(np.float32(2.0) * np.float32(self.half_map_height))

    def to_c(self) -> MbSettings:
        s = MbSettings()
        s.enable_simulation = int(bool(self.enable_simulation))
        s.enable_spring = int(bool(self.enable_spring))
        s.number_of_grids_one_dimension = int(self.number_of_grids_one_dimension)
        s.grid_size = int(self.grid_size)
        s.half_map_height = int(self.half_map_height)
        for i in range(3):
            s.gravity[i] = float(self.gravity[i])
        s.e_envir = float(self.e_envir)
        s.e_agent = float(self.e_agent)
        s.k_agent = float(self.k_agent)
        s.c_agent = float(self.c_agent)
        s.r_agent = float(self.r_agent)
        s.max_number_of_agents = int(self._max_number_of_agents)
        s.time_delta = float(self.time_delta)
        return s

    @classmethod
    def from_c(cls, s: MbSettings) -> "PhysicsSettings":
        return cls(
            enable_simulation=bool(s.enable_simulation),
            enable_spring=bool(s.enable_spring),
            number_of_grids_one_dimension=int(s.number_of_grids_one_dimension),
            grid_size=int(s.grid_size),
            half_map_height=int(s.half_map_height),
            gravity=(float(s.gravity[0]), float(s.gravity[1]), float(s.gravity[2])),
            e_envir=float(s.e_envir),
            e_agent=float(s.e_agent),
            k_agent=float(s.k_agent),
            c_agent=float(s.c_agent),
            r_agent=float(s.r_agent),
            time_delta=float(s.time_delta),
            _max_number_of_agents=int(s.max_number_of_agents),
        )


def inspect_checkpoint(data: bytes) -> Tuple[PhysicsSettings, int, int]:
    """(settings, time, number of spheres) of a checkpoint; needs no device."""
    buf = C.create_string_buffer(bytes(data), len(data))
    s = MbSettings()
    t = C.c_int64()
    n = C.c_uint32()
    check(lib().mb_checkpoint_inspect(buf, len(data), C.byref(s), C.byref(t), C.byref(n)))
    return PhysicsSettings.from_c(s), t.value, n.value


def default_settings_from_library() -> PhysicsSettings:
    s = MbSettings()
    lib().mb_settings_default(C.byref(s))
    return PhysicsSettings.from_c(s)


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _stream_ptr(stream) -> Optional[int]:
    if stream is None:
        return None
    return int(getattr(stream, "cuda_stream", stream))


class Physics:
    """One simulated world: the ping-pong buffers, the frame counter and the
    per-frame sort -> spatial hashing -> physics sequence."""

    def __init__(self, settings: PhysicsSettings):
        self._settings = settings
        self._h = C.c_void_p()
        cs = settings.to_c()
        check(lib().mb_physics_create(C.byref(cs), C.byref(self._h)))
        self.number_of_agents = settings.max_number_of_agents()

    # -- lifetime -------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib().mb_physics_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self) -> "Physics":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # -- settings (ResMut<PhysicsSettings>) ------------------------------
    @property
    def settings(self) -> PhysicsSettings:
        s = MbSettings()
        check(lib().mb_physics_get_settings(self._h, C.byref(s)))
        return PhysicsSettings.from_c(s)

    def set_settings(self, settings: PhysicsSettings) -> None:
        cs = settings.to_c()
        check(lib().mb_physics_set_settings(self._h, C.byref(cs)))
        self._settings = settings

    # -- state ------------------------------------------------------------
    def upload_state(self, position_with_scale: np.ndarray, packed_vfme: np.ndarray,
                     stream=None) -> None:
        pos = np.ascontiguousarray(position_with_scale, dtype=np.float32).reshape(-1, 4)
        vf = np.ascontiguousarray(packed_vfme, dtype=np.uint32).reshape(-1, 4)
        if pos.shape[0] != self.number_of_agents or vf.shape[0] != self.number_of_agents:
            raise ValueError(f"state must hold {self.number_of_agents} spheres")
        check(lib().mb_physics_upload_state(self._h, _ptr(pos), _ptr(vf), _stream_ptr(stream)))

    def upload_state_device(self, position_with_scale, packed_vfme, stream=None) -> None:
        """Device tensors (N x 4 float32, N x 4 int32/uint32) already on the GPU."""
        check(lib().mb_physics_upload_state_device(
            self._h, C.c_void_p(position_with_scale.data_ptr()),
            C.c_void_p(packed_vfme.data_ptr()), _stream_ptr(stream)))

    def download_state(self, stream=None) -> Tuple[np.ndarray, np.ndarray]:
        pos = np.empty((self.number_of_agents, 4), np.float32)
        vf = np.empty((self.number_of_agents, 4), np.uint32)
        check(lib().mb_physics_download_state(self._h, _ptr(pos), _ptr(vf), _stream_ptr(stream)))
        return pos, vf

    # -- tuning (launch choices that do not change results) ---------------
    @property
    def tuning(self) -> Tuple[float, int, int, int, int]:
        """(scan_margin, stage_pool, row_sort_threads, dense_scan, one_launch)."""
        t = MbTuning()
        check(lib().mb_physics_get_tuning(self._h, C.byref(t)))
        return (float(t.scan_margin), int(t.stage_pool), int(t.row_sort_threads),
                int(t.dense_scan), int(t.one_launch))

    def set_tuning(self, scan_margin: float, stage_pool: int, row_sort_threads: int = 512,
                   dense_scan: int = 0, one_launch: int = 0) -> None:
        t = MbTuning()
        check(lib().mb_physics_get_tuning(self._h, C.byref(t)))
        t.scan_margin, t.stage_pool = float(scan_margin), int(stage_pool)
        t.row_sort_threads, t.dense_scan = int(row_sort_threads), int(dense_scan)
        t.one_launch = int(one_launch)
        check(lib().mb_physics_set_tuning(self._h, C.byref(t)))

    def set_wait_limit(self, polls: int) -> None:
        """Polls of one bounded wait of the one-launch frame (0: the default);
        for tests of its failure path."""
        check(lib().mb_physics_debug_set_wait_limit(self._h, int(polls)))

    @property
    def frame_graph(self) -> bool:
        """Whether runs of LOOP frames replay a captured two-frame hipGraph."""
        t = MbTuning()
        check(lib().mb_physics_get_tuning(self._h, C.byref(t)))
        return bool(t.frame_graph)

    def set_frame_graph(self, enable: bool) -> None:
        t = MbTuning()
        check(lib().mb_physics_get_tuning(self._h, C.byref(t)))
        t.frame_graph = int(bool(enable))
        check(lib().mb_physics_set_tuning(self._h, C.byref(t)))

    def autotune(self, frames_per_trial: int = 5) -> Tuple[Tuple[float, int, int, int, int], float]:
        """Times every candidate tuning on a copy of the state, keeps the
        fastest; returns ((scan_margin, stage_pool, row_sort_threads,
        dense_scan, one_launch), ms per frame)."""
        t = MbTuning()
        ms = C.c_double()
        check(lib().mb_physics_autotune(self._h, int(frames_per_trial), C.byref(t), C.byref(ms)))
        return (float(t.scan_margin), int(t.stage_pool), int(t.row_sort_threads),
                int(t.dense_scan), int(t.one_launch)), ms.value

    # -- checkpoints --------------------------------------------------------
    def save_checkpoint(self) -> bytes:
        """Settings, time and the last frame's state (see mb_physics_save_checkpoint)."""
        size = lib().mb_checkpoint_size(self.number_of_agents)
        buf = (C.c_char * size)()
        check(lib().mb_physics_save_checkpoint(self._h, buf, size))
        return bytes(buf)

    def load_checkpoint(self, data: bytes) -> None:
        buf = C.create_string_buffer(bytes(data), len(data))
        check(lib().mb_physics_load_checkpoint(self._h, buf, len(data)))
        self._settings = self.settings

    # -- stepping (PhysicsNode::update + run) ----------------------------
    def run(self, frames: int = 1, stream=None) -> None:
        check(lib().mb_physics_step(self._h, int(frames), _stream_ptr(stream)))

    def time(self) -> int:
        return int(lib().mb_physics_time(self._h))

    def is_even(self) -> Optional[bool]:
        r = lib().mb_physics_is_even(self._h)
        return None if r < 0 else bool(r)

    def buffers(self) -> MbBuffers:
        b = MbBuffers()
        check(lib().mb_physics_buffers(self._h, C.byref(b)))
        return b

    def debug(self, stream=None) -> dict:
        n = self.number_of_agents
        out = {k: np.empty(n, np.uint32) for k in ("perm", "start", "close", "next_keys")}
        check(lib().mb_physics_debug_download(
            self._h, _ptr(out["perm"]), _ptr(out["start"]), _ptr(out["close"]),
            _ptr(out["next_keys"]), _stream_ptr(stream)))
        return out

    def set_profiling(self, enable: bool) -> None:
        check(lib().mb_physics_set_profiling(self._h, int(bool(enable))))

    def profile_read(self, capacity: int = 1 << 16) -> np.ndarray:
        """(frames, 2) device milliseconds of [sort (with the cell ranges),
        collide + integrate (with the next frame's bucket filing)] per
        profiled frame since the last read."""
        buf = np.empty((capacity, 2), np.float64)
        f = C.c_uint32()
        check(lib().mb_physics_profile_read(self._h, _ptr(buf), capacity, C.byref(f)))
        return buf[: f.value].copy()


# ---------------------------------------------------------------------------
# Stage entry points on device tensors (bevy_radix_sort::run,
# spatial_hashing::run, physics::run)
# ---------------------------------------------------------------------------

def radix_sort_workspace_size(max_number_of_keys: int) -> int:
    return int(lib().mb_radix_sort_workspace_size(int(max_number_of_keys)))


def radix_sort_run(workspace, keys_in, vals_in, keys_out, vals_out, number_of_keys: int,
                   pass_range: range, stream=None) -> None:
    """``bevy_radix_sort::run(.., number_of_keys, pass_range, init_vals, ..)``.

    ``vals_in=None`` is ``init_vals = true``.  Arguments are device tensors.
    """
    check(lib().mb_radix_sort(
        C.c_void_p(workspace.data_ptr()), C.c_void_p(keys_in.data_ptr()),
        None if vals_in is None else C.c_void_p(vals_in.data_ptr()),
        C.c_void_p(keys_out.data_ptr()), C.c_void_p(vals_out.data_ptr()),
        int(number_of_keys), int(pass_range.start), int(pass_range.stop), _stream_ptr(stream)))


def spatial_hashing_run(sorted_keys, start, close, number_of_keys: int, zero_count=None,
                        stream=None) -> None:
    """``spatial_hashing::run(.., number_of_keys, ..)`` on device tensors."""
    zc = number_of_keys if zero_count is None else zero_count
    check(lib().mb_spatial_hashing(
        C.c_void_p(sorted_keys.data_ptr()), C.c_void_p(start.data_ptr()),
        C.c_void_p(close.data_ptr()), int(number_of_keys), int(zc), _stream_ptr(stream)))


def physics_run(settings: PhysicsSettings, init: bool, perm, start, close, key_out, pos_i,
                vfme_i, pos_o, vfme_o, number_of_agents: int, stream=None) -> None:
    """``physics::run`` with one parity's bind group spelled out."""
    cs = settings.to_c()

    def p(t):
        return None if t is None else C.c_void_p(t.data_ptr())

    check(lib().mb_physics_run(
        C.byref(cs), int(bool(init)), p(perm), p(start), p(close), p(key_out), p(pos_i),
        p(vfme_i), p(pos_o), p(vfme_o), int(number_of_agents), _stream_ptr(stream)))


def check_exact_rewrites(div_samples: int = 1 << 28):
    """Runs the library's self-check of the contact solve's exact rewrites on
    the current device: (sqrt mismatches over every float in [2^-96, 2^96],
    quotient mismatches over `div_samples` operand triples); both 0 on a
    correct build."""
    s = C.c_uint64(0)
    d = C.c_uint64(0)
    check(lib().mb_debug_check_rewrites(int(div_samples), C.byref(s), C.byref(d)))
    return int(s.value), int(d.value)


__all__ = [
    "PhysicsPlugin",
    "PhysicsSettings",
    "Physics",
    "MbError",
    "MB_INVALID_INDEX",
    "NUMBER_OF_THREADS_PER_WORKGROUP",
    "default_settings_from_library",
    "radix_sort_workspace_size",
    "radix_sort_run",
    "check_exact_rewrites",
    "spatial_hashing_run",
    "physics_run",
]
