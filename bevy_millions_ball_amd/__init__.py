"""MI355X-native physics step of bevy_millions_ball.

The compute path is libmillions_ball.so (hand-written HIP for gfx950) behind
the C ABI in include/millions_ball.h; this package is the host-side mirror of
the reference's PhysicsPlugin surface over that ABI.
"""
from .physics import (  # noqa: F401
    MB_INVALID_INDEX,
    NUMBER_OF_THREADS_PER_WORKGROUP,
    MbError,
    Physics,
    PhysicsPlugin,
    PhysicsSettings,
    default_settings_from_library,
    inspect_checkpoint,
    physics_run,
    radix_sort_run,
    radix_sort_workspace_size,
    spatial_hashing_run,
)
from .scenes import CONFIGS, SceneConfig, generate, pack_vfme, unpack_vfme  # noqa: F401
