"""ctypes front end of oracle/libmb_oracle.so (TEST INFRASTRUCTURE ONLY).

Pinned by (tests/test_oracle_*.py):
  f16 pack/unpack  - exhaustively against the F16C instruction and numpy;
  radix sort       - the reference's spatial_hashing.rs:551-585 known answers
                     (keys i % 256) and numpy's stable argsort;
  spatial hashing  - the same known answers (start/close closed forms);
  quantisation     - closed-form f32 values of the shader constants;
  physics          - closed forms (free fall, wall bounce, head-on restitution,
                     overlap push, spring force, hash values) and invariants;
                     the neighbour walk by closed-form pairs across each of the
                     eight neighbour directions and the grid's edges, and by the
                     visited candidate list against a brute-force restatement of
                     the 3 x 3 neighbourhood (tests/test_oracle_walk.py), which
                     mutated copies of the walk fail (tests/test_oracle_mutants.py);
  cell key         - the reference's unclamped key, on the demo lattice of
                     main.rs:42-49 and on points outside the map.
Physics at the level of the adapter's own rounding is parity unpinned: only a
wgpu run of physics.wgsl could fix it, and none is available.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmb_oracle.so"
# Another build of the same source may be named here (the mutation tests load
# deliberately broken copies of mb_oracle.c to show that the walk checks fail).
LIB_OVERRIDE = os.environ.get("MB_ORACLE_LIB")


class OSettings(C.Structure):
    _fields_ = [
        ("enable_simulation", C.c_int32),
        ("enable_spring", C.c_int32),
        ("number_of_grids_one_dimension", C.c_uint32),
        ("grid_size", C.c_uint32),
        ("half_map_height", C.c_uint32),
        ("gravity", C.c_float * 3),
        ("e_envir", C.c_float),
        ("e_agent", C.c_float),
        ("k_agent", C.c_float),
        ("c_agent", C.c_float),
        ("r_agent", C.c_float),
        ("max_number_of_agents", C.c_uint32),
        ("time_delta", C.c_float),
    ]


class OConsts(C.Structure):
    _fields_ = [
        ("grids", C.c_uint32),
        ("grid_size", C.c_uint32),
        ("half_map_height_u", C.c_uint32),
        ("half_map_size_u", C.c_uint32),
        ("half_map_height", C.c_float),
        ("half_map_size", C.c_float),
        ("gravity", C.c_float * 3),
        ("e_envir", C.c_float),
        ("e_agent", C.c_float),
        ("k_agent", C.c_float),
        ("c_agent", C.c_float),
        ("r_agent", C.c_float),
        ("t_delta", C.c_float),
        ("enable_spring", C.c_int32),
    ]


_lib = None
_P = C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if LIB_OVERRIDE:
            L = C.CDLL(LIB_OVERRIDE)
        else:
            if not LIB_PATH.exists():
                build()
            L = C.CDLL(str(LIB_PATH))
        L.mbo_settings_default.argtypes = [C.POINTER(OSettings)]
        L.mbo_compile_consts.argtypes = [C.POINTER(OSettings), C.POINTER(OConsts)]
        L.mbo_f32_to_f16.argtypes = [C.c_float]
        L.mbo_f32_to_f16.restype = C.c_uint16
        L.mbo_f16_to_f32.argtypes = [C.c_uint16]
        L.mbo_f16_to_f32.restype = C.c_float
        L.mbo_check_f16_against_f16c.argtypes = [C.c_int]
        L.mbo_check_f16_against_f16c.restype = C.c_int64
        L.mbo_hash3f1u.argtypes = [C.c_uint32, C.POINTER(C.c_float)]
        L.mbo_cell_key.argtypes = [C.POINTER(OConsts), C.c_float, C.c_float]
        L.mbo_cell_key.restype = C.c_uint32
        L.mbo_radix_sort.argtypes = [_P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
        L.mbo_spatial_hashing.argtypes = [_P, _P, _P, C.c_uint32, C.c_uint32, C.c_int]
        L.mbo_physics_run.argtypes = [C.POINTER(OConsts), C.c_int, _P, _P, _P, C.c_uint32, _P,
                                      _P, _P, _P, _P, C.c_uint32, C.c_int]
        L.mbo_physics_candidates.argtypes = [C.POINTER(OConsts), _P, _P, _P, C.c_uint32, _P, _P,
                                             C.c_uint32, C.c_uint32, _P, C.c_uint32]
        L.mbo_physics_candidates.restype = C.c_uint32
        L.mbo_world_create.argtypes = [C.POINTER(OSettings)]
        L.mbo_world_create.restype = _P
        L.mbo_world_destroy.argtypes = [_P]
        L.mbo_world_set_consts.argtypes = [_P, C.POINTER(OSettings)]
        L.mbo_world_upload.argtypes = [_P, _P, _P]
        L.mbo_world_step.argtypes = [_P, C.c_uint32, C.c_int, C.POINTER(C.c_double)]
        L.mbo_world_time.argtypes = [_P]
        L.mbo_world_time.restype = C.c_int64
        L.mbo_world_download.argtypes = [_P, _P, _P]
        L.mbo_world_debug.argtypes = [_P, _P, _P, _P]
        L.mbo_world_next_keys.argtypes = [_P, _P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def default_threads() -> int:
    for var in ("MB_ORACLE_THREADS", "OMP_NUM_THREADS"):
        if os.environ.get(var):
            return max(1, int(os.environ[var]))
    return os.cpu_count() or 1


def settings_from(ps) -> OSettings:
    """OSettings from a PhysicsSettings-like object (field names of the
    reference's PhysicsSettings)."""
    s = OSettings()
    s.enable_simulation = int(bool(ps.enable_simulation))
    s.enable_spring = int(bool(ps.enable_spring))
    s.number_of_grids_one_dimension = int(ps.number_of_grids_one_dimension)
    s.grid_size = int(ps.grid_size)
    s.half_map_height = int(ps.half_map_height)
    for i in range(3):
        s.gravity[i] = float(ps.gravity[i])
    s.e_envir = float(ps.e_envir)
    s.e_agent = float(ps.e_agent)
    s.k_agent = float(ps.k_agent)
    s.c_agent = float(ps.c_agent)
    s.r_agent = float(ps.r_agent)
    mna = ps.max_number_of_agents
    s.max_number_of_agents = int(mna() if callable(mna) else mna)
    s.time_delta = float(ps.time_delta)
    return s


def compile_consts(ps) -> OConsts:
    c = OConsts()
    s = settings_from(ps)
    lib().mbo_compile_consts(C.byref(s), C.byref(c))
    return c


def f32_to_f16(x: float) -> int:
    return int(lib().mbo_f32_to_f16(float(x)))


def f16_to_f32(h: int) -> float:
    return float(lib().mbo_f16_to_f32(int(h)))


def hash3f1u(n: int) -> np.ndarray:
    out = (C.c_float * 3)()
    lib().mbo_hash3f1u(int(n), out)
    return np.array(out[:], np.float32)


def cell_key(consts: OConsts, x: float, z: float) -> int:
    return int(lib().mbo_cell_key(C.byref(consts), float(x), float(z)))


def radix_sort(keys: np.ndarray, vals=None, pass_range=range(0, 3), threads=None):
    keys = np.ascontiguousarray(keys, np.uint32)
    n = keys.shape[0]
    ko = np.empty(n, np.uint32)
    vo = np.empty(n, np.uint32)
    v = None if vals is None else np.ascontiguousarray(vals, np.uint32)
    lib().mbo_radix_sort(_p(keys), _p(v), _p(ko), _p(vo), n, pass_range.start, pass_range.stop,
                         threads or default_threads())
    return ko, vo


def spatial_hashing(sorted_keys: np.ndarray, start=None, close=None, zero_count=None,
                    threads=None):
    k = np.ascontiguousarray(sorted_keys, np.uint32)
    n = k.shape[0]
    zc = n if zero_count is None else zero_count
    size = max(zc, n, 1)
    st = np.zeros(size, np.uint32) if start is None else start
    cl = np.zeros(size, np.uint32) if close is None else close
    lib().mbo_spatial_hashing(_p(k), _p(st), _p(cl), n, zc, threads or default_threads())
    return st, cl


def physics_run(ps, init, perm, start, close, pos_i, vfme_i, threads=None):
    c = compile_consts(ps)
    n = perm.shape[0]
    key_out = np.empty(n, np.uint32)
    pos_o = np.empty((n, 4), np.float32)
    vfme_o = np.empty((n, 4), np.uint32)
    cells = 0 if start is None else int(min(len(start), len(close)))
    lib().mbo_physics_run(C.byref(c), int(bool(init)), _p(np.ascontiguousarray(perm, np.uint32)),
                          _p(start), _p(close), cells, _p(key_out),
                          _p(np.ascontiguousarray(pos_i, np.float32)),
                          _p(np.ascontiguousarray(vfme_i, np.uint32)), _p(pos_o), _p(vfme_o), n,
                          threads or default_threads())
    return key_out, pos_o, vfme_o


def physics_candidates(ps, perm, start, close, pos_i, vfme_i, slot, cap=4096):
    """Sorted slots the LOOP pass of `slot` tests against, in its order."""
    c = compile_consts(ps)
    perm = np.ascontiguousarray(perm, np.uint32)
    start = np.ascontiguousarray(start, np.uint32)
    close = np.ascontiguousarray(close, np.uint32)
    pos_i = np.ascontiguousarray(pos_i, np.float32)
    vfme_i = np.ascontiguousarray(vfme_i, np.uint32)
    out = np.zeros(cap, np.uint32)
    cnt = lib().mbo_physics_candidates(C.byref(c), _p(perm), _p(start), _p(close),
                                       int(min(len(start), len(close))), _p(pos_i), _p(vfme_i),
                                       perm.shape[0], int(slot), _p(out), cap)
    return out[:min(int(cnt), cap)].copy()


class World:
    """The reference's per-frame order over its ping-pong buffers, on the CPU."""

    def __init__(self, ps):
        self._s = settings_from(ps)
        self.n = int(self._s.max_number_of_agents)
        self._w = lib().mbo_world_create(C.byref(self._s))

    def close(self):
        if self._w:
            lib().mbo_world_destroy(self._w)
            self._w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_settings(self, ps):
        self._s = settings_from(ps)
        lib().mbo_world_set_consts(self._w, C.byref(self._s))

    def upload(self, pos, vfme):
        pos = np.ascontiguousarray(pos, np.float32)
        vfme = np.ascontiguousarray(vfme, np.uint32)
        lib().mbo_world_upload(self._w, _p(pos), _p(vfme))

    def step(self, frames=1, threads=None, stage_seconds=None):
        buf = (C.c_double * 3)()
        lib().mbo_world_step(self._w, int(frames), threads or default_threads(), buf)
        if stage_seconds is not None:
            for i in range(3):
                stage_seconds[i] += buf[i]

    def time(self) -> int:
        return int(lib().mbo_world_time(self._w))

    def download(self):
        pos = np.empty((self.n, 4), np.float32)
        vf = np.empty((self.n, 4), np.uint32)
        lib().mbo_world_download(self._w, _p(pos), _p(vf))
        return pos, vf

    def debug(self):
        g = int(self._s.number_of_grids_one_dimension)
        cells = max(self.n, g * g)
        out = {k: np.empty(self.n, np.uint32) for k in ("perm", "next_keys")}
        out.update({k: np.empty(cells, np.uint32) for k in ("start", "close")})
        lib().mbo_world_debug(self._w, _p(out["perm"]), _p(out["start"]), _p(out["close"]))
        lib().mbo_world_next_keys(self._w, _p(out["next_keys"]))
        return out


def check_f16_against_f16c(threads=None) -> int:
    return int(lib().mbo_check_f16_against_f16c(threads or default_threads()))
