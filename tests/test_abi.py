"""The C ABI library loads, exports every function include/millions_ball.h
declares, and validates settings before touching a device (CPU-only)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from bevy_millions_ball_amd import _lib
from bevy_millions_ball_amd.physics import PhysicsPlugin, PhysicsSettings

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "millions_ball.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_struct_layout():
    assert C.sizeof(_lib.MbSettings) == 60
    assert C.sizeof(_lib.MbBuffers) == 8 * 8 + 8


def test_defaults_match_reference():
    # PhysicsPlugin::default, src/physics/mod.rs:87-109
    s = _lib.MbSettings()
    _lib.lib().mb_settings_default(C.byref(s))
    want = PhysicsSettings.from_plugin(PhysicsPlugin())
    assert bytes(s) == bytes(want.to_c())  # field by field, as f32
    assert (want.number_of_grids_one_dimension, want.grid_size, want.half_map_height) == (128, 8, 16)
    assert want.max_number_of_agents() == 16 * 1024
    assert want.enable_simulation and not want.enable_spring


def test_settings_accessors():
    # mod.rs:364-400
    ps = PhysicsSettings(number_of_grids_one_dimension=64, grid_size=8, half_map_height=8,
                         _max_number_of_agents=256 * 256)
    assert ps.max_number_of_grids_on_dimension() == 256
    assert ps.map_size_u() == 512 and ps.half_map_size_u() == 256
    assert ps.map_size() == 512.0 and ps.half_map_size() == 256.0
    assert ps.half_map_height_f32() == 8.0 and ps.map_height() == 16.0


@pytest.mark.parametrize("field,value,fragment", [
    ("number_of_grids_one_dimension", 0, "grids"),
    ("grid_size", 0, "grid_size"),
    ("number_of_grids_one_dimension", 200, "exceeds max_number_of_agents"),
    ("number_of_grids_one_dimension", 8192, "2^24"),
    ("max_number_of_agents", 0, "max_number_of_agents"),
])
def test_create_rejects_invalid_settings(field, value, fragment):
    s = PhysicsSettings(number_of_grids_one_dimension=128, grid_size=1,
                        _max_number_of_agents=128 * 128)
    cs = s.to_c()
    setattr(cs, field, value)
    if field == "number_of_grids_one_dimension" and value == 8192:
        cs.max_number_of_agents = (1 << 30) - 1
    h = C.c_void_p()
    rc = _lib.lib().mb_physics_create(C.byref(cs), C.byref(h))
    assert rc == _lib.MB_ERR_INVALID_ARGUMENT
    assert not h.value
    assert fragment in _lib.lib().mb_last_error().decode()


def test_stage_entry_points_validate_arguments():
    L = _lib.lib()
    assert L.mb_radix_sort(None, None, None, None, None, 10, 0, 3, None) == _lib.MB_ERR_INVALID_ARGUMENT
    assert L.mb_radix_sort(C.c_void_p(16), C.c_void_p(32), None, C.c_void_p(64), C.c_void_p(128),
                           10, 0, 5, None) == _lib.MB_ERR_INVALID_ARGUMENT
    assert L.mb_radix_sort(None, None, None, None, None, 0, 0, 3, None) == _lib.MB_OK  # nothing to do
    assert L.mb_radix_sort_workspace_size(1 << 20) >= (1 << 20) * 8
    assert L.mb_physics_run(None, 0, None, None, None, None, None, None, None, None, 4, None) \
        == _lib.MB_ERR_INVALID_ARGUMENT


def test_package_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()
