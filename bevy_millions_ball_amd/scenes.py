"""Seeded synthetic scenes and the per-sphere record format.

The record is the reference's: ``position_with_scale`` float4 (xyz, w =
radius) and ``packed_vfme`` uint4 of f16 pairs (vx|vy, vz|fx, fy|fz,
mass|0), packed like ``pack_vfme`` in src/physics/mod.rs:214-239 (Rust
``as f16`` rounds to nearest even, as numpy's float16 cast does).  Velocity is
a per-step displacement (physics.wgsl:255-266), so "velocities U[-1,1]^3" are
in those native units.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np

from .physics import PhysicsSettings


def pack_vfme(velocity: np.ndarray, force: np.ndarray, mass: np.ndarray) -> np.ndarray:
    """(N,3) f32 velocity, (N,3) f32 force, (N,) f32 mass -> (N,4) uint32."""
    v = np.asarray(velocity, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    f = np.asarray(force, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    m = np.asarray(mass, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    out = np.empty((v.shape[0], 4), np.uint32)
    out[:, 0] = v[:, 0] | (v[:, 1] << 16)
    out[:, 1] = v[:, 2] | (f[:, 0] << 16)
    out[:, 2] = f[:, 1] | (f[:, 2] << 16)
    out[:, 3] = m
    return out


def unpack_vfme(packed: np.ndarray):
    """(N,4) uint32 -> velocity (N,3), force (N,3), mass (N,), all float32."""
    p = np.asarray(packed, np.uint32)
    lo = (p & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float32)
    hi = (p >> 16).astype(np.uint16).view(np.float16).astype(np.float32)
    velocity = np.stack([lo[:, 0], hi[:, 0], lo[:, 1]], axis=1)
    force = np.stack([hi[:, 1], lo[:, 2], hi[:, 2]], axis=1)
    return velocity, force, lo[:, 3].copy()


@dataclass(frozen=True)
class SceneConfig:
    """One benchmark workload: the physics settings plus how spheres are drawn."""

    name: str
    number_of_agents: int
    grids: int
    half_map_height: int = 8
    grid_size: int = 1
    radius: float = 0.5
    time_delta: float = 1.0 / 60.0
    e_agent: float = 0.8
    e_envir: float = 0.8
    enable_spring: bool = False
    k_agent: float = 25.0
    c_agent: float = 0.1
    r_agent: float = 0.9
    layout: str = "uniform"  # or "clumps"
    clumps: int = 64
    clump_sigma_cells: float = 8.0
    steps: int = 100

    def settings(self) -> PhysicsSettings:
        return PhysicsSettings(
            enable_simulation=True,
            enable_spring=self.enable_spring,
            number_of_grids_one_dimension=self.grids,
            grid_size=self.grid_size,
            half_map_height=self.half_map_height,
            gravity=(0.0, -9.8, 0.0),
            e_envir=self.e_envir,
            e_agent=self.e_agent,
            k_agent=self.k_agent,
            c_agent=self.c_agent,
            r_agent=self.r_agent,
            time_delta=self.time_delta,
            _max_number_of_agents=self.number_of_agents,
        )


# The five benchmark workloads.  The spring workload leaves r_agent and the
# restitution unspecified; the PhysicsPlugin default r_agent 0.9 and e = 0.8
# are used.
CONFIGS = {
    "plain_65k": SceneConfig("plain_65k", 65_536, 256, steps=100),
    "uniform_1m": SceneConfig("uniform_1m", 1_048_576, 1024, steps=1000),
    "spring_clumps_1m": SceneConfig(
        "spring_clumps_1m", 1_048_576, 1024, time_delta=1.0 / 120.0, enable_spring=True,
        k_agent=500.0, c_agent=5.0, layout="clumps", steps=1000),
    "uniform_4m": SceneConfig("uniform_4m", 4_194_304, 2048, steps=500),
}


def generate(cfg: SceneConfig, seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """Seeded (numpy PCG64) initial state for `cfg`, strictly inside the box."""
    rng = np.random.default_rng(seed)
    n = cfg.number_of_agents
    r = np.float32(cfg.radius)
    hms = np.float32((cfg.grids * cfg.grid_size) // 2)
    hmh = np.float32(cfg.half_map_height)
    lo_xz, hi_xz = -hms + r, hms - r
    pos = np.empty((n, 4), np.float32)
    if cfg.layout == "uniform":
        pos[:, 0] = rng.uniform(lo_xz, hi_xz, n)
        pos[:, 2] = rng.uniform(lo_xz, hi_xz, n)
    elif cfg.layout == "clumps":
        centres = rng.uniform(lo_xz + 3 * cfg.clump_sigma_cells, hi_xz - 3 * cfg.clump_sigma_cells,
                              (cfg.clumps, 2))
        which = rng.integers(0, cfg.clumps, n)
        sigma = cfg.clump_sigma_cells * cfg.grid_size
        xz = centres[which] + rng.normal(0.0, sigma, (n, 2))
        xz = np.clip(xz, lo_xz, hi_xz)
        pos[:, 0] = xz[:, 0]
        pos[:, 2] = xz[:, 1]
    else:
        raise ValueError(cfg.layout)
    pos[:, 1] = rng.uniform(r, np.float32(2.0) * hmh - r, n)
    pos[:, 3] = r
    vel = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    force = np.zeros((n, 3), np.float32)
    mass = np.ones(n, np.float32)
    return pos, pack_vfme(vel, force, mass)


def reference_lattice(max_number_of_agents: int, seed: int = 0):
    """The demo's start state (mod.rs:172-212): an XZ lattice with spacing 4 at
    y = 4, radius 1, velocities U[-0.8, 0.8) in x and z.  The reference draws
    the velocities with ChaCha8 seeded 0; numpy's PCG64 is used here, so the
    lattice is exact and the velocities are only distributed alike."""
    n = max_number_of_agents
    log = int(n).bit_length() - 1
    log_x = log // 2
    log_z = log - log_x
    cx, cz = 1 << log_x, 1 << log_z
    xs = (np.arange(cx, dtype=np.float32) - np.float32(0.5) * (np.float32(cx) - np.float32(1.0)))
    zs = (np.arange(cz, dtype=np.float32) - np.float32(0.5) * (np.float32(cz) - np.float32(1.0)))
    gx, gz = np.meshgrid(xs, zs, indexing="ij")
    pos = np.empty((cx * cz, 4), np.float32)
    pos[:, 0] = gx.reshape(-1) * np.float32(4.0)
    pos[:, 1] = 4.0
    pos[:, 2] = gz.reshape(-1) * np.float32(4.0)
    pos[:, 3] = 1.0
    rng = np.random.default_rng(seed)
    vel = np.zeros((n, 3), np.float32)
    vel[:, 0] = rng.uniform(-0.8, 0.8, n)
    vel[:, 2] = rng.uniform(-0.8, 0.8, n)
    return pos[:n], pack_vfme(vel, np.zeros((n, 3), np.float32), np.ones(n, np.float32))
