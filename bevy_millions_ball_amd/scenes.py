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


# ---- the demo's seeded velocity stream -------------------------------------
# mod.rs:199-212 draws each sphere's velocity with
# `ChaCha8Rng::seed_from_u64(0)` and `rng.gen_range(-0.8..0.8)` twice (x, then
# z).  Those crates are not in the reference tree; their published algorithms
# at the versions Cargo.lock pins are restated here:
#   rand_core 0.6  SeedableRng::seed_from_u64: PCG32 (XSH-RR; multiplier
#                  6364136223846793005, increment 11634580027462260723; the
#                  state is advanced before each output) fills the 32-byte
#                  seed four little-endian bytes at a time;
#   rand_chacha 0.3.1  ChaCha8Rng: the ChaCha block function with 8 rounds,
#                  key = seed, a 64-bit block counter from 0 in words 12-13 and
#                  stream 0 in words 14-15; next_u32 reads the keystream as
#                  little-endian words, block after block;
#   rand 0.8.5     UniformFloat<f32>::sample_single(low, high): draw a u32,
#                  keep its top 23 bits as the mantissa of a float in [1, 2),
#                  subtract 1, then res = value * (high - low) + low (two f32
#                  roundings, no fused multiply-add); redraw while res >= high.

_CHACHA_CONST = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574], np.uint32)
_PCG_MUL = 6364136223846793005
_PCG_INC = 11634580027462260723
_M64 = (1 << 64) - 1


def pcg_xsh_rr(state: int) -> int:
    """PCG's XSH-RR output function: 32 bits from a 64-bit state."""
    xorshifted = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
    rot = state >> 59
    return ((xorshifted >> rot) | (xorshifted << ((32 - rot) & 31))) & 0xFFFFFFFF


def pcg32_seed_bytes(state: int, length: int = 32) -> bytes:
    """rand_core 0.6 SeedableRng::seed_from_u64's seed bytes."""
    out = bytearray()
    state &= _M64
    while len(out) < length:
        state = (state * _PCG_MUL + _PCG_INC) & _M64
        out += pcg_xsh_rr(state).to_bytes(4, "little")
    return bytes(out[:length])


def _rotl(v, c):
    return (v << np.uint32(c)) | (v >> np.uint32(32 - c))


def chacha_blocks(key: bytes, first_block: int, count: int, rounds: int = 8,
                  stream: int = 0) -> np.ndarray:
    """`count` consecutive ChaCha blocks (16 little-endian u32 words each,
    block counter first_block, first_block + 1, ...), flattened."""
    kw = np.frombuffer(key, dtype="<u4").astype(np.uint32)
    ctr = np.arange(first_block, first_block + count, dtype=np.uint64)
    init = np.empty((16, count), np.uint32)
    init[0:4] = _CHACHA_CONST[:, None]
    init[4:12] = kw[:, None]
    init[12] = (ctr & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    init[13] = (ctr >> np.uint64(32)).astype(np.uint32)
    init[14] = np.uint32(stream & 0xFFFFFFFF)
    init[15] = np.uint32((stream >> 32) & 0xFFFFFFFF)
    x = [init[i].copy() for i in range(16)]

    def qr(a, b, c, d):
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 16)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 12)
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 8)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 7)

    for _ in range(rounds // 2):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    out = np.stack([x[i] + init[i] for i in range(16)])  # (16, count)
    return out.T.reshape(-1)


class ChaCha8U32:
    """rand_chacha 0.3.1 ChaCha8Rng::seed_from_u64(seed) as a u32 stream."""

    def __init__(self, seed: int, rounds: int = 8):
        self.key = pcg32_seed_bytes(seed)
        self.rounds = rounds
        self.block = 0
        self.buf = np.empty(0, np.uint32)
        self.pos = 0

    def take(self, count: int) -> np.ndarray:
        while self.buf.shape[0] - self.pos < count:
            need = count - (self.buf.shape[0] - self.pos)
            blocks = max(4, (need + 15) // 16)
            more = chacha_blocks(self.key, self.block, blocks, self.rounds)
            self.block += blocks
            self.buf = np.concatenate([self.buf[self.pos:], more])
            self.pos = 0
        out = self.buf[self.pos:self.pos + count]
        self.pos += count
        return out


def gen_range_f32(words: np.ndarray, low: float, high: float) -> np.ndarray:
    """rand 0.8.5 UniformFloat<f32>::sample_single on successive u32 draws,
    without the redraw (candidates; the caller drops those >= high)."""
    lo = np.float32(low)
    hi = np.float32(high)
    scale = np.float32(hi - lo)
    value1_2 = ((words >> np.uint32(9)) | np.uint32(127 << 23)).view(np.float32)
    value0_1 = value1_2 - np.float32(1.0)
    return (value0_1 * scale).astype(np.float32) + lo


def demo_velocities(n: int, seed: int = 0) -> np.ndarray:
    """Velocities of packed_vfme_data (mod.rs:199-212): x and z drawn in that
    order per sphere from ChaCha8 seeded `seed`, U[-0.8, 0.8); y = 0."""
    rng = ChaCha8U32(seed)
    need = 2 * n
    out = np.empty(need, np.float32)
    got = 0
    while got < need:
        cand = gen_range_f32(rng.take(need - got), -0.8, 0.8)
        ok = cand < np.float32(0.8)
        good = cand.shape[0] if ok.all() else int(np.argmin(ok))
        out[got:got + good] = cand[:good]
        got += good
        if good < cand.shape[0]:
            # the rejected draw is consumed; the words after it are drawn again
            rng.pos -= cand.shape[0] - good - 1
    vel = np.zeros((n, 3), np.float32)
    vel[:, 0] = out[0::2]
    vel[:, 2] = out[1::2]
    return vel


def reference_lattice(max_number_of_agents: int, seed: int = 0):
    """The demo's start state (mod.rs:172-212): an XZ lattice with spacing 4 at
    y = 4, radius 1, velocities U[-0.8, 0.8) in x and z from ChaCha8 seeded 0
    (demo_velocities), mass 1, force 0."""
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
    vel = demo_velocities(n, seed)
    return pos[:n], pack_vfme(vel, np.zeros((n, 3), np.float32), np.ones(n, np.float32))
