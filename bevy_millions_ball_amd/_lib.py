"""ctypes binding of libmillions_ball.so (include/millions_ball.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmillions_ball.so"

MB_OK = 0
MB_ERR_INVALID_ARGUMENT = -1
MB_ERR_HIP = -2
MB_ERR_OUT_OF_MEMORY = -3
MB_ERR_STATE = -4
MB_INVALID_INDEX = 0xFFFFFFFF


class MbSettings(C.Structure):
    """``mb_settings``: field for field the reference's PhysicsPlugin."""

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


class MbBuffers(C.Structure):
    _fields_ = [
        ("position_with_scale", C.c_void_p * 2),
        ("packed_vfme", C.c_void_p * 2),
        ("global_keys", C.c_void_p * 2),
        ("global_vals", C.c_void_p * 2),
        ("number_of_agents", C.c_uint32),
    ]


class MbTuning(C.Structure):
    _fields_ = [("scan_margin", C.c_float), ("stage_pool", C.c_uint32),
                ("row_sort_threads", C.c_uint32), ("frame_graph", C.c_uint32),
                ("dense_scan", C.c_uint32), ("one_launch", C.c_uint32)]


class MbServerConfig(C.Structure):
    _fields_ = [
        ("memory_cap_bytes", C.c_uint64),
        ("chunk_bytes", C.c_uint64),
        ("use_graph", C.c_int32),
        ("profile_frames", C.c_int32),
        ("admission", C.c_int32),
        ("reserved", C.c_int32),
    ]


MB_ADMIT_FIFO = 0
MB_ADMIT_MOST_STEPS_FIRST = 1


class MbRequestStatus(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("number_of_agents", C.c_uint32),
        ("steps", C.c_uint32),
        ("frames_done", C.c_uint32),
        ("submit_s", C.c_double),
        ("admit_s", C.c_double),
        ("done_s", C.c_double),
    ]


class MbServerStats(C.Structure):
    _fields_ = [
        ("queued", C.c_uint32),
        ("running", C.c_uint32),
        ("retiring", C.c_uint32),
        ("done", C.c_uint32),
        ("used_bytes", C.c_uint64),
        ("cap_bytes", C.c_uint64),
        ("reserved_bytes", C.c_uint64),
        ("frames", C.c_uint64),
        ("sphere_steps", C.c_double),
    ]


MB_REQUEST_QUEUED = 0
MB_REQUEST_RUNNING = 1
MB_REQUEST_RETIRING = 2
MB_REQUEST_DONE = 3


class MbError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"millions_ball error {code}: {message}")
        self.code = code


_lib = None

_P = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
_F32P = C.POINTER(C.c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "mb_settings_default": (None, [C.POINTER(MbSettings)]),
    "mb_physics_create": (C.c_int, [C.POINTER(MbSettings), C.POINTER(_P)]),
    "mb_physics_destroy": (C.c_int, [_P]),
    "mb_physics_set_settings": (C.c_int, [_P, C.POINTER(MbSettings)]),
    "mb_physics_get_settings": (C.c_int, [_P, C.POINTER(MbSettings)]),
    "mb_physics_upload_state": (C.c_int, [_P, _P, _P, _P]),
    "mb_physics_upload_state_device": (C.c_int, [_P, _P, _P, _P]),
    "mb_physics_download_state": (C.c_int, [_P, _P, _P, _P]),
    "mb_physics_step": (C.c_int, [_P, C.c_uint32, _P]),
    "mb_physics_time": (C.c_int64, [_P]),
    "mb_physics_is_even": (C.c_int, [_P]),
    "mb_physics_buffers": (C.c_int, [_P, C.POINTER(MbBuffers)]),
    "mb_physics_debug_download": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "mb_physics_set_profiling": (C.c_int, [_P, C.c_int]),
    "mb_physics_profile_read": (C.c_int, [_P, _P, C.c_uint32, _U32P]),
    "mb_physics_debug_set_wait_limit": (C.c_int, [_P, C.c_uint32]),
    "mb_last_error": (C.c_char_p, []),
    "mb_radix_sort_workspace_size": (C.c_size_t, [C.c_uint32]),
    "mb_radix_sort": (
        C.c_int,
        [_P, _P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P],
    ),
    "mb_radix_sort_static": (
        C.c_int,
        [_P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P],
    ),
    "mb_debug_check_rewrites": (
        C.c_int, [C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mb_spatial_hashing": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, _P]),
    "mb_physics_run": (
        C.c_int,
        [C.POINTER(MbSettings), C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, C.c_uint32, _P],
    ),
    "mb_physics_get_tuning": (C.c_int, [_P, C.POINTER(MbTuning)]),
    "mb_physics_set_tuning": (C.c_int, [_P, C.POINTER(MbTuning)]),
    "mb_physics_autotune": (
        C.c_int, [_P, C.c_uint32, C.POINTER(MbTuning), C.POINTER(C.c_double)]),
    "mb_checkpoint_size": (C.c_size_t, [C.c_uint32]),
    "mb_physics_save_checkpoint": (C.c_int, [_P, _P, C.c_size_t]),
    "mb_physics_load_checkpoint": (C.c_int, [_P, _P, C.c_size_t]),
    "mb_checkpoint_inspect": (
        C.c_int, [_P, C.c_size_t, C.POINTER(MbSettings), C.POINTER(C.c_int64), _U32P]),
    "mb_server_scene_bytes": (C.c_uint64, [C.c_uint32, C.c_uint32]),
    "mb_server_create": (C.c_int, [C.POINTER(MbServerConfig), C.POINTER(_P)]),
    "mb_server_destroy": (C.c_int, [_P]),
    "mb_server_submit": (
        C.c_int,
        [_P, C.POINTER(MbSettings), _P, _P, C.c_uint32, C.c_int, C.POINTER(C.c_uint64)],
    ),
    "mb_server_tick": (C.c_int, [_P, _U32P]),
    "mb_server_drain": (C.c_int, [_P]),
    "mb_server_status": (C.c_int, [_P, C.c_uint64, C.POINTER(MbRequestStatus)]),
    "mb_server_collect": (C.c_int, [_P, C.c_uint64, _P, _P, _P, _P]),
    "mb_server_stats_read": (C.c_int, [_P, C.POINTER(MbServerStats)]),
    "mb_server_frame_times": (C.c_int, [_P, _P, C.c_uint32, _U32P]),
}


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m bevy_millions_ball_amd.build`"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != MB_OK:
        raise MbError(rc, lib().mb_last_error().decode(errors="replace"))
