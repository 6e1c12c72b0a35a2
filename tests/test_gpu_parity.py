"""GPU parity: the HIP path (through the C ABI) against the oracle.

Integer stages (cell keys, sorted permutation, start/close) must be bit-exact.
State is compared bit-exactly too: the kernels follow the oracle's operation
order, so any difference is a bug (the stated bound is 1e-5 relative for f32
positions and unpacked f16 velocity/force; `assert_state_close` reports it
when bit-equality fails so the size of a mismatch is visible)."""
import json
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from bevy_millions_ball_amd import Physics, PhysicsSettings, radix_sort_run, \
    radix_sort_workspace_size, spatial_hashing_run, physics_run
from bevy_millions_ball_amd.scenes import CONFIGS, SceneConfig, generate, pack_vfme, unpack_vfme

pytestmark = pytest.mark.gpu

KAT = json.loads((Path(__file__).parent / "golden" / "spatial_hashing_kat.json").read_text())
INVALID = 0xFFFFFFFF


def dev(a, device):
    import torch

    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to(device)


def host_u32(t):
    return t.cpu().numpy().view(np.uint32)


def assert_state_close(pos_a, vf_a, pos_b, vf_b, rtol=1e-5, atol=1e-6):
    if np.array_equal(pos_a.view(np.uint32), pos_b.view(np.uint32)) and np.array_equal(vf_a, vf_b):
        return
    va, fa, ma = unpack_vfme(vf_a)
    vb, fb, mb = unpack_vfme(vf_b)
    for name, x, y in (("pos", pos_a, pos_b), ("vel", va, vb), ("force", fa, fb), ("mass", ma, mb)):
        bad = ~np.isclose(x, y, rtol=rtol, atol=atol, equal_nan=True)
        assert not bad.any(), f"{name}: {bad.sum()} mismatches, first {np.argwhere(bad)[:3].tolist()}"
    raise AssertionError("state differs at bit level (within 1e-5 relative)")


# ---------------------------------------------------------------------------
# Stage entry points
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", KAT["cases"],
                         ids=lambda c: f"n{c['number_of_keys']}-p{c['pass_count']}-{c['read_from_even']}")
def test_sort_and_hashing_reference_known_answers(gpu, case):
    import torch

    n = case["number_of_keys"]
    print(f"[kat] n={n} passes={case['pass_count']}", flush=True)
    keys = dev(np.arange(n, dtype=np.uint32) % 256, gpu)
    vals = dev(np.arange(n, dtype=np.uint32), gpu)
    ko = torch.empty_like(keys)
    vo = torch.empty_like(keys)
    ws = torch.zeros(radix_sort_workspace_size(n), dtype=torch.uint8, device=gpu)
    radix_sort_run(ws, keys, vals, ko, vo, n, range(0, case["pass_count"]))
    start = torch.empty(max(n, 1), dtype=torch.int32, device=gpu)
    close = torch.empty_like(start)
    spatial_hashing_run(ko, start, close, n)
    cells = len(case["start"])
    assert host_u32(start)[:cells].tolist() == case["start"]
    assert host_u32(close)[:cells].tolist() == case["close"]
    perm = host_u32(vo)
    assert np.array_equal(perm, np.argsort(np.arange(n) % 256, kind="stable").astype(np.uint32))


@pytest.mark.parametrize("n", [1, 2, 63, 64, 255, 4095, 4096, 4097, 12_345, 100_000, 1_048_576])
@pytest.mark.parametrize("passes", [1, 3, 4])
def test_sort_random_matches_stable_argsort(gpu, n, passes):
    import torch

    rng = np.random.default_rng(n * 7 + passes)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    ws = torch.zeros(radix_sort_workspace_size(n), dtype=torch.uint8, device=gpu)
    kd = dev(keys, gpu)
    vd = dev(vals, gpu)
    ko = torch.empty_like(kd)
    vo = torch.empty_like(kd)
    mask = np.uint32(0xFFFFFFFF if passes == 4 else (1 << (8 * passes)) - 1)
    order = np.argsort(keys & mask, kind="stable")
    for rep in range(2):  # the workspace is reused: it must clean up after itself
        radix_sort_run(ws, kd, vd, ko, vo, n, range(0, passes))
        assert np.array_equal(host_u32(ko), keys[order])
        assert np.array_equal(host_u32(vo), vals[order])
    radix_sort_run(ws, kd, None, ko, vo, n, range(0, passes))  # init_vals = true
    assert np.array_equal(host_u32(vo), order.astype(np.uint32))


def test_sort_nearly_sorted_skewed_keys(gpu):
    """Keys like a physics frame: nearly sorted, long runs of equal high bytes."""
    import torch

    rng = np.random.default_rng(9)
    n = 300_000
    keys = np.sort(rng.integers(0, 1 << 20, n)).astype(np.uint32)
    swap = rng.integers(0, n - 1, n // 10)
    keys[swap], keys[swap + 1] = keys[swap + 1], keys[swap]
    keys[rng.integers(0, n, 1000)] = 0
    ws = torch.zeros(radix_sort_workspace_size(n), dtype=torch.uint8, device=gpu)
    kd = dev(keys, gpu)
    ko = torch.empty_like(kd)
    vo = torch.empty_like(kd)
    radix_sort_run(ws, kd, None, ko, vo, n, range(0, 3))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(host_u32(vo), order.astype(np.uint32))


def test_spatial_hashing_random(gpu, oracle):
    import torch

    rng = np.random.default_rng(4)
    n = 200_000
    keys = np.sort(rng.integers(0, 150_000, n)).astype(np.uint32)
    start = torch.zeros(n, dtype=torch.int32, device=gpu)
    close = torch.zeros(n, dtype=torch.int32, device=gpu)
    spatial_hashing_run(dev(keys, gpu), start, close, n)
    s_ref, c_ref = oracle.spatial_hashing(keys, start=np.zeros(n, np.uint32), close=np.zeros(n, np.uint32))
    s = host_u32(start)
    c = host_u32(close)
    assert np.array_equal(s, s_ref)
    valid = s_ref != INVALID
    assert np.array_equal(c[valid], c_ref[valid])


@pytest.mark.parametrize("spring", [False, True])
def test_physics_run_stage_matches_oracle(gpu, oracle, spring):
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=8192, grids=64, enable_spring=spring,
                  k_agent=500.0, c_agent=5.0)
    ps = cfg.settings()
    pos, vf = generate(cfg, seed=21)
    n = cfg.number_of_agents
    # a sort + hash of the INIT keys gives the LOOP inputs
    c = oracle.compile_consts(ps)
    keys = np.array([oracle.cell_key(c, x, z) for x, z in zip(pos[:, 0], pos[:, 2])], np.uint32)
    sk, perm = oracle.radix_sort(keys, None, range(0, 3))
    start, close = oracle.spatial_hashing(sk, start=np.zeros(n, np.uint32), close=np.zeros(n, np.uint32))
    for init in (True, False):
        k_ref, p_ref, v_ref = oracle.physics_run(ps, init, perm, start, close, pos, vf)
        import torch

        ko = torch.empty(n, dtype=torch.int32, device=gpu)
        po = torch.empty((n, 4), dtype=torch.float32, device=gpu)
        vo = torch.empty((n, 4), dtype=torch.int32, device=gpu)
        physics_run(ps, init, dev(perm, gpu), dev(start, gpu), dev(close, gpu), ko,
                    dev(pos, gpu), dev(vf, gpu), po, vo, n)
        assert np.array_equal(host_u32(ko), k_ref)
        assert_state_close(po.cpu().numpy(), host_u32(vo).reshape(n, 4), p_ref, v_ref)


# ---------------------------------------------------------------------------
# Whole frames through mb_physics_step
# ---------------------------------------------------------------------------

def compare_frames(gpu_world, ref_world, frames, check_every=1):
    for f in range(frames):
        gpu_world.run(1)
        ref_world.step(1)
        assert gpu_world.time() == ref_world.time()
        if (f + 1) % check_every and f != frames - 1:
            continue
        d = gpu_world.debug()
        r = ref_world.debug()
        assert np.array_equal(d["perm"], r["perm"]), f"perm differs at frame {gpu_world.time()}"
        assert np.array_equal(d["start"], r["start"]), f"start differs at frame {gpu_world.time()}"
        valid = r["start"] != INVALID
        cells = np.nonzero(valid)[0]
        assert np.array_equal(d["close"][cells], r["close"][cells])
        assert np.array_equal(d["next_keys"], r["next_keys"])
        pg, vg = gpu_world.download_state()
        pr, vr = ref_world.download()
        assert_state_close(pg, vg, pr, vr)


def make_pair(oracle, cfg: SceneConfig, seed):
    ps = cfg.settings()
    pos, vf = generate(cfg, seed=seed)
    g = Physics(ps)
    g.upload_state(pos, vf)
    r = oracle.World(ps)
    r.upload(pos, vf)
    return g, r


@pytest.mark.parametrize("name,n,grids,frames", [
    ("plain_65k", 4096, 64, 40),
    ("plain_65k", 65_536, 256, 12),
    ("spring_clumps_1m", 16_384, 128, 40),
    ("uniform_1m", 20_000, 128, 20),   # N not a multiple of the tile
])
def test_step_bit_exact_small(gpu, oracle, name, n, grids, frames):
    cfg = replace(CONFIGS[name], number_of_agents=n, grids=grids)
    print(f"[step] {name} n={n} G={grids} frames={frames}", flush=True)
    g, r = make_pair(oracle, cfg, seed=n)
    with g:
        compare_frames(g, r, frames)


@pytest.mark.parametrize("name,frames", [("uniform_1m", 3), ("spring_clumps_1m", 3),
                                         ("uniform_4m", 2)])
def test_step_bit_exact_full_size(gpu, oracle, name, frames):
    cfg = CONFIGS[name]
    print(f"[full] {name} frames={frames}", flush=True)
    g, r = make_pair(oracle, cfg, seed=1)
    with g:
        compare_frames(g, r, frames)


@pytest.mark.parametrize("name,frames", [("uniform_1m", 200), ("spring_clumps_1m", 60),
                                         ("uniform_4m", 40)])
def test_long_run_bit_exact_full_size(gpu, oracle, name, frames):
    """Hundreds of frames at full size, compared every 20 frames: the
    trajectories are chaotic, so a single differing bit anywhere (a contact
    order, a rounding, a filing) would spread and show."""
    cfg = CONFIGS[name]
    print(f"[long] {name} frames={frames}", flush=True)
    g, r = make_pair(oracle, cfg, seed=7)
    with g:
        compare_frames(g, r, frames, check_every=20)


def test_long_run_dense_scan_full_size(gpu, oracle):
    """The spring-clumps scene at the point the autotuner picks for it (margin
    1/16, 1408-slot pool, dense scan: height-slab filtered walks, and in the
    first frames many rescans after contacts that move a sphere beyond the
    margin), 60 frames at full size, compared every 20."""
    cfg = CONFIGS["spring_clumps_1m"]
    print("[long-dense] spring_clumps_1m frames=60", flush=True)
    g, r = make_pair(oracle, cfg, seed=7)
    with g:
        g.set_tuning(0.0625, 1408, 512, 1)
        compare_frames(g, r, 60, check_every=20)


@pytest.mark.parametrize("spring", [False, True])
@pytest.mark.parametrize("half_height", [1, 8, 64])
def test_dense_scan_mixed_radii_and_heights(gpu, oracle, spring, half_height):
    """The dense-scan walk tests only the candidates whose height slab (16
    slabs over the box height) meets the lane's reach, which comes from the
    block's largest radius.  Clumps of up to ~90 spheres per cell (up to
    ~700 candidates a sphere) with radii from 0.1 to 0.9, in boxes 2, 16 and
    128 tall (slabs far thinner than the reach, about as tall, far taller),
    must give the oracle's frames, with the 1408-slot pool (nearly every block
    staged) and with the 896-slot one (most blocks overflow it and walk from
    global memory)."""
    rng = np.random.default_rng(31 + half_height + (7 if spring else 0))
    n, grids, gs = 16_384, 64, 2
    ps = PhysicsSettings(number_of_grids_one_dimension=grids, grid_size=gs,
                         half_map_height=half_height, enable_spring=spring, e_agent=0.8,
                         e_envir=0.8, time_delta=1 / 120, _max_number_of_agents=n)
    hms = grids * gs // 2
    centres = rng.uniform(-hms / 2, hms / 2, (4, 2))
    xz = centres[rng.integers(0, 4, n)] + rng.normal(0.0, 8.0, (n, 2))
    xz = np.clip(xz, -hms + 1.0, hms - 1.0)
    r = rng.uniform(0.1, 0.9, n).astype(np.float32)
    pos = np.empty((n, 4), np.float32)
    pos[:, 0], pos[:, 2] = xz[:, 0], xz[:, 1]
    pos[:, 1] = rng.uniform(r, 2 * half_height - r)
    pos[:, 3] = r
    vf = pack_vfme(rng.uniform(-1, 1, (n, 3)).astype(np.float32), np.zeros((n, 3), np.float32),
                   rng.uniform(0.5, 2.0, n).astype(np.float32))
    ref = oracle.World(ps)
    ref.upload(pos, vf)
    ref.step(6)
    want = ref.download(), ref.debug()["next_keys"]
    ref.close()
    for tuning in [(0.0625, 1408, 512, 1), (1.0, 896, 512, 1)]:
        print(f"[dense-radii] spring={spring} H={half_height} {tuning}", flush=True)
        with Physics(ps) as g:
            g.set_tuning(*tuning)
            g.upload_state(pos, vf)
            g.run(6)
            pg, vg = g.download_state()
            assert np.array_equal(pg.view(np.uint32), want[0][0].view(np.uint32)), tuning
            assert np.array_equal(vg, want[0][1]), tuning
            assert np.array_equal(g.debug()["next_keys"], want[1]), tuning


def test_largest_grid_bit_exact(gpu, oracle):
    """The largest grid the keys allow (G = 4096, G^2 = 2^24 cells) with
    N = G^2 spheres: ~4096 spheres per row, about half of them above the
    common row sort's 4096 in-LDS places, which makes the mean row (4096 >
    3072) pick the variant with 8192 places."""
    cfg = replace(CONFIGS["uniform_1m"], number_of_agents=1 << 24, grids=4096)
    print("[max] G=4096 n=16777216", flush=True)
    g, r = make_pair(oracle, cfg, seed=3)
    with g:
        compare_frames(g, r, 3)
        rows = np.bincount(r.debug()["next_keys"] // 4096, minlength=4096)
        assert rows.max() > 4096 > rows.min()


def test_hundred_frames_statistics(gpu, oracle):
    """Over 100 frames the trajectories stay bit-identical; the kinetic energy
    and penetration statistics are then trivially equal, and are checked for
    being finite and bounded as a sanity property."""
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=16_384, grids=128)
    g, r = make_pair(oracle, cfg, seed=100)
    with g:
        g.run(101)
        r.step(101)
        pg, vg = g.download_state()
        pr, vr = r.download()
        assert_state_close(pg, vg, pr, vr)
        v, _, m = unpack_vfme(vg)
        ke = 0.5 * float(np.sum(m[:, None] * v.astype(np.float64) ** 2))
        assert np.isfinite(ke) and ke < 16_384 * 3


def test_time_parity_and_buffers(gpu, oracle):
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=4096, grids=64)
    g, r = make_pair(oracle, cfg, seed=2)
    with g:
        assert g.time() == -1 and g.is_even() is None
        g.run(1)
        assert g.time() == 0 and g.is_even() is True
        g.run(1)
        assert g.time() == 1 and g.is_even() is False
        b = g.buffers()
        assert b.number_of_agents == 4096
        assert len({b.position_with_scale[0], b.position_with_scale[1], b.packed_vfme[0],
                    b.packed_vfme[1]}) == 4


def test_settings_change_and_disable(gpu, oracle):
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=4096, grids=64)
    g, r = make_pair(oracle, cfg, seed=3)
    with g:
        compare_frames(g, r, 3)
        ps = cfg.settings()
        ps.enable_spring = True
        ps.k_agent, ps.c_agent = 500.0, 5.0
        ps.e_envir = 0.5
        g.set_settings(ps)
        r.set_settings(ps)
        compare_frames(g, r, 3)
        ps.enable_simulation = False
        g.set_settings(ps)
        t = g.time()
        g.run(5)
        assert g.time() == t


@pytest.mark.parametrize("split,grids,grid_size", [
    (1, 32, 2),    # same map, coarser cells; right after INIT
    (4, 32, 2),
    (3, 32, 1),    # a smaller map: the state lies partly outside it (keys alias later rows)
    (5, 90, 1),    # a larger map with finer cells, G not a power of two
])
def test_regrid_matches_the_reference_frames(gpu, oracle, split, grids, grid_size):
    """A live grid change (the slider of main.rs:190-200, mod.rs:808-837
    recompiles the pipelines and carries on with the buffers as they are):
    the frame after it sorts the keys the last frame wrote under the old
    grid and walks start / close with the new grid's cells, stale close words
    included; the frames after it are ordinary.  Frame by frame against the
    oracle, which keeps the reference's buffers."""
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=8192, grids=64)
    print(f"[regrid] split={split} -> G={grids} gs={grid_size}", flush=True)
    g, r = make_pair(oracle, cfg, seed=11)
    with g:
        compare_frames(g, r, split)
        ps2 = replace(cfg.settings(), number_of_grids_one_dimension=grids, grid_size=grid_size)
        g.set_settings(ps2)
        r.set_settings(ps2)
        compare_frames(g, r, 6)
    r.close()


def demo_settings():
    """main.rs:42-49 over the PhysicsPlugin defaults (mod.rs:87-109)."""
    return PhysicsSettings(number_of_grids_one_dimension=64, half_map_height=8, e_agent=0.8,
                           e_envir=0.8, _max_number_of_agents=256 * 256)


def test_reference_demo_scene_bit_exact(gpu, oracle):
    """The reference demo: 65,536 spheres on a 64 x 64 grid of cells of size 8
    (a +-256 map) starting from the lattice of mod.rs:173-196, which reaches
    |x| = 510, with the ChaCha8 velocities of mod.rs:199-212.  INIT keys cells
    past the grid (u32(x + 256) / 8 up to 95: keys alias later rows, up to
    6175); INIT and the frame that sorts those keys run on the stage kernels
    in the reference's buffer roles, then the frame path takes over."""
    from bevy_millions_ball_amd.scenes import reference_lattice

    ps = demo_settings()
    pos, vf = reference_lattice(65536)
    g = Physics(ps)
    g.upload_state(pos, vf)
    r = oracle.World(ps)
    r.upload(pos, vf)
    with g:
        compare_frames(g, r, 1)
        keys = g.debug()["next_keys"]
        assert keys.max() == 6175 and (keys >= 64 * 64).sum() > 0
        compare_frames(g, r, 40)
    r.close()


def test_upload_rejects_walks_past_the_arrays(gpu):
    """A lattice of 4096 spheres reaching |x| = 126 in a +-32 map: the walk of
    a sphere at x = 126 reads cell keys past N = 4096, where the reference
    reads start / close out of bounds.  The upload is refused, and stepping
    without an accepted state is an error."""
    from bevy_millions_ball_amd.scenes import reference_lattice

    n = 64 * 64
    ps = PhysicsSettings(number_of_grids_one_dimension=64, grid_size=1, half_map_height=8,
                         e_agent=0.8, e_envir=0.8, _max_number_of_agents=n)
    pos, vf = reference_lattice(n)
    with Physics(ps) as g:
        with pytest.raises(Exception, match="out of bounds"):
            g.upload_state(pos, vf)
        with pytest.raises(Exception, match="no accepted state"):
            g.run(1)
        # negative radii are refused too
        cfg = replace(CONFIGS["plain_65k"], number_of_agents=n, grids=64)
        good, gvf = generate(cfg, seed=1)
        bad = good.copy()
        bad[9, 3] = -0.5
        with pytest.raises(Exception, match="negative radius"):
            g.upload_state(bad, gvf)
        # an in-map state is accepted afterwards
        g.upload_state(good, gvf)
        g.run(2)


@pytest.mark.parametrize("mode", ["spring_r_agent_1", "tiny_radius"])
def test_outputs_that_can_leave_the_grid(gpu, oracle, mode):
    """States whose LOOP outputs can reach past the map's +x / +z edge: spring
    walls with r_agent = 1 (the top of the demo's slider, main.rs:186) clamp
    a sphere at HALF_MAP_SIZE itself, and a radius far below the f32 spacing
    at the edge does the same with rigid walls.  Every frame then runs on the
    stage kernels; keys >= G^2 appear and must match the oracle's."""
    rng = np.random.default_rng(31)
    G, n = 64, 8192
    hms = G // 2
    if mode == "spring_r_agent_1":
        ps = PhysicsSettings(enable_spring=True, number_of_grids_one_dimension=G, grid_size=1,
                             half_map_height=8, k_agent=500.0, c_agent=5.0, r_agent=1.0,
                             e_agent=0.8, e_envir=0.8, time_delta=1 / 120, _max_number_of_agents=n)
        radius = 0.5
    else:
        ps = PhysicsSettings(number_of_grids_one_dimension=G, grid_size=1, half_map_height=8,
                             e_agent=0.8, e_envir=0.8, time_delta=1 / 60, _max_number_of_agents=n)
        radius = 1e-7
    pos = np.empty((n, 4), np.float32)
    pos[:, 0] = rng.uniform(-hms + 0.5, hms - 0.5, n)
    pos[:, 2] = rng.uniform(-hms + 0.5, hms - 0.5, n)
    pos[:, 1] = rng.uniform(0.5, 15.5, n)
    pos[:, 3] = radius
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    vf = pack_vfme(vel, np.zeros((n, 3), np.float32), np.ones(n, np.float32))
    g = Physics(ps)
    g.upload_state(pos, vf)
    r = oracle.World(ps)
    r.upload(pos, vf)
    past = 0
    with g:
        for _ in range(12):
            compare_frames(g, r, 1)
            past += int((g.debug()["next_keys"] >= G * G).sum())
    r.close()
    assert past > 0


def test_checkpoint_of_an_out_of_grid_state(gpu):
    """A checkpoint of the demo's INIT frame holds a state outside the grid:
    loading it keys the state as the reference would (unclamped) and the
    next frame sorts those keys on the stage kernels, so resuming equals the
    uninterrupted run."""
    from bevy_millions_ball_amd.scenes import reference_lattice

    ps = demo_settings()
    pos, vf = reference_lattice(65536)
    with Physics(ps) as a:
        a.upload_state(pos, vf)
        a.run(1)
        ck = a.save_checkpoint()
        a.run(5)
        want = a.download_state(), a.debug()["next_keys"]
    with Physics(ps) as b:
        b.load_checkpoint(ck)
        b.run(5)
        got = b.download_state(), b.debug()["next_keys"]
    assert np.array_equal(got[0][0].view(np.uint32), want[0][0].view(np.uint32))
    assert np.array_equal(got[0][1], want[0][1])
    assert np.array_equal(got[1], want[1])


def checkpoint_state(ck: bytes, n: int):
    head = 4 + 4 + 60 + 8 + 4
    pos = np.frombuffer(ck, np.float32, n * 4, head).reshape(n, 4)
    vf = np.frombuffer(ck, np.uint32, n * 4, head + n * 16).reshape(n, 4)
    return pos, vf


def test_synchronous_calls_wait_for_the_callers_stream(gpu, oracle):
    """Frames enqueued on a side stream (a torch stream does not synchronise
    with the null stream), then at once a checkpoint, a grid change and an
    autotune: each must act on the state those frames leave, not on an
    earlier one."""
    import torch

    cfg = CONFIGS["uniform_1m"]  # ~50 us per frame: a missing wait would show
    n = cfg.number_of_agents
    side = torch.cuda.Stream()
    g, r = make_pair(oracle, cfg, seed=4)
    with g:
        g.run(4, stream=side)
        ck = g.save_checkpoint()
        r.step(4)
        pr, vr = r.download()
        pc, vc = checkpoint_state(ck, n)
        assert np.array_equal(pc.view(np.uint32), pr.view(np.uint32)) and np.array_equal(vc, vr)
        g.run(3, stream=side)
        r.step(3)
        ps2 = replace(cfg.settings(), number_of_grids_one_dimension=512, grid_size=2)
        g.set_settings(ps2)
        r.set_settings(ps2)
        g.run(2, stream=side)
        r.step(2)
        g.autotune(1)
        side.synchronize()
        pg, vg = g.download_state()
        pr, vr = r.download()
        assert_state_close(pg, vg, pr, vr)
        assert np.array_equal(g.debug()["next_keys"], r.debug()["next_keys"])
    r.close()


@pytest.mark.parametrize("name,n,grids", [("plain_65k", 65_536, 256), ("spring_clumps_1m", 65_536, 256)])
def test_frame_graph_replays_match_the_oracle(gpu, oracle, name, n, grids):
    """Runs of LOOP frames replay a captured two-frame hipGraph.  Replays are
    interleaved with eager work (downloads, profiled frames launched directly,
    an odd-length run that starts on an odd frame) and every replay is
    checked against the oracle, then against a handle that never uses the
    graph."""
    cfg = replace(CONFIGS[name], number_of_agents=n, grids=grids)
    print(f"[graph] {name} n={n}", flush=True)
    g, r = make_pair(oracle, cfg, seed=17)
    h, _ = make_pair(oracle, cfg, seed=17)
    with g, h:
        assert not g.frame_graph  # off by default (slower than direct launches here)
        g.set_frame_graph(True)
        assert g.frame_graph and not h.frame_graph
        for run in (2, 4, 3, 6, 1, 8):
            g.run(run)
            h.run(run)
            r.step(run)
            pg, vg = g.download_state()
            pr, vr = r.download()
            assert_state_close(pg, vg, pr, vr)
            assert np.array_equal(g.debug()["next_keys"], r.debug()["next_keys"])
            if run == 6:
                g.set_profiling(True)
                g.run(2)
                h.run(2)
                r.step(2)
                assert g.profile_read().shape == (2, 2)
                g.set_profiling(False)
        ph, vh = h.download_state()
        pg, vg = g.download_state()
        assert np.array_equal(pg.view(np.uint32), ph.view(np.uint32)) and np.array_equal(vg, vh)
    r.close()


def test_edge_cases_tiny_worlds(gpu, oracle):
    rng = np.random.default_rng(8)
    # one sphere, one cell
    for n, grids in ((1, 1), (64, 1), (256, 16), (257, 16)):
        ps = PhysicsSettings(number_of_grids_one_dimension=grids, grid_size=2, half_map_height=4,
                             e_agent=0.8, e_envir=0.8, time_delta=1 / 60,
                             _max_number_of_agents=n)
        hms = grids  # G * 2 / 2
        pos = np.empty((n, 4), np.float32)
        pos[:, 0] = rng.uniform(-hms + 0.5, hms - 0.5, n)
        pos[:, 2] = rng.uniform(-hms + 0.5, hms - 0.5, n)
        pos[:, 1] = rng.uniform(0.5, 7.5, n)
        pos[:, 3] = 0.5
        vf = pack_vfme(rng.uniform(-1, 1, (n, 3)).astype(np.float32), np.zeros((n, 3), np.float32),
                       np.ones(n, np.float32))
        g = Physics(ps)
        g.upload_state(pos, vf)
        r = oracle.World(ps)
        r.upload(pos, vf)
        with g:
            compare_frames(g, r, 8)


def test_exact_rewrites_against_ieee(gpu):
    """The contact solve's short square root equals sqrtf on every float in
    [2^-96, 2^96], and its shared-reciprocal quotients equal '/' on 2^28
    operand triples (divisors inside and outside the rewrite's range; zero,
    signed-zero, +-d, tiny and slightly-above-d numerators), bit for bit, on
    the device itself."""
    from bevy_millions_ball_amd.physics import check_exact_rewrites

    print("[rewrites] sqrt exhaustive, div 2^28", flush=True)
    sqrt_bad, div_bad = check_exact_rewrites(1 << 28)
    assert sqrt_bad == 0, f"{sqrt_bad} square roots differ from sqrtf"
    assert div_bad == 0, f"{div_bad} quotient triples differ from '/'"


def test_exact_rewrite_fallback_paths(gpu, oracle):
    """The kernels' exact rewrites next to their general code, in one wave:
    contacts whose offset has a component below 2^-60 (general division),
    an offset with a signed zero (shared reciprocal), and radii at and above
    the half map size (compare-and-select wall clamp with -e >= e)."""
    rng = np.random.default_rng(21)
    n, grids = 48, 4
    ps = PhysicsSettings(number_of_grids_one_dimension=grids, grid_size=1, half_map_height=4,
                         e_agent=0.8, e_envir=0.8, time_delta=1 / 60,
                         _max_number_of_agents=n)
    pos = np.empty((n, 4), np.float32)
    pos[:, 0] = rng.uniform(-1.5, 1.5, n)
    pos[:, 2] = rng.uniform(-1.5, 1.5, n)
    pos[:, 1] = rng.uniform(0.5, 7.5, n)
    pos[:, 3] = 0.5
    pos[0] = (1e-30, 2.0, 0.25, 0.5)     # against sphere 1: |t.x| = 1e-30 < 2^-60
    pos[1] = (0.0, 2.6, 0.25, 0.5)
    pos[2] = (-0.0, 5.0, -1.0, 0.5)      # against sphere 3: t.x = -0
    pos[3] = (0.0, 5.5, -1.2, 0.5)
    pos[4] = (0.3, 3.0, 0.3, 2.5)        # radius above the half map size (2)
    pos[5] = (-0.7, 6.0, 0.9, 2.0)       # radius equal to it: bounds of zero
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    vel[:6] = 0.0
    vf = pack_vfme(vel, np.zeros((n, 3), np.float32), np.ones(n, np.float32))
    g = Physics(ps)
    g.upload_state(pos, vf)
    r = oracle.World(ps)
    r.upload(pos, vf)
    with g:
        compare_frames(g, r, 6)


def test_dense_rows_spill_and_fallbacks(gpu, oracle):
    """Rows far above the mean: bucket spill, the global-memory row sort and
    the collide path that reads neighbours from global memory."""
    from scene_helpers import dense_row_scene

    ps, pos, vf = dense_row_scene()
    g = Physics(ps)
    g.upload_state(pos, vf)
    r = oracle.World(ps)
    r.upload(pos, vf)
    with g:
        compare_frames(g, r, 5)
        # the spill path was taken: rows still far above the bucket capacity
        keys = r.debug()["next_keys"]
        rows = np.bincount(keys // ps.number_of_grids_one_dimension, minlength=64)
        assert rows.max() > 1024


@pytest.mark.parametrize("split", [0, 1, 6])
def test_checkpoint_resume_is_bit_exact(gpu, split):
    """Save after `split` frames (0: before INIT), resume in a fresh handle
    and in the same handle after stepping elsewhere: both continue exactly
    like the uninterrupted run."""
    from bevy_millions_ball_amd import inspect_checkpoint

    cfg = replace(CONFIGS["plain_65k"], number_of_agents=8192, grids=64)
    ps = cfg.settings()
    pos, vf = generate(cfg, seed=5)
    with Physics(ps) as a:
        a.upload_state(pos, vf)
        a.run(split)
        ck = a.save_checkpoint()
        assert inspect_checkpoint(ck)[1:] == (split - 1, 8192)
        a.run(7)
        want = a.download_state(), a.debug()["next_keys"]
        a.load_checkpoint(ck)  # rewind the same handle
        assert a.time() == split - 1
        a.run(7)
        again = a.download_state(), a.debug()["next_keys"]
    with Physics(ps) as b:
        b.load_checkpoint(ck)
        b.run(7)
        got = b.download_state(), b.debug()["next_keys"]
    for res in (again, got):
        assert np.array_equal(res[0][0].view(np.uint32), want[0][0].view(np.uint32))
        assert np.array_equal(res[0][1], want[0][1])
        assert np.array_equal(res[1], want[1])


TUNINGS = [(m, p, s, d, o) for o in (0, 1) for d in (0, 1) for s in (512, 256)
           for p in (896, 1408) for m in (1.0, 0.5, 0.25, 0.0625) if not (o and s != 256)]


@pytest.mark.parametrize("scene", ["uniform", "dense", "clumps"])
def test_every_tuning_matches_the_oracle(gpu, oracle, scene):
    """Every point the autotuner can pick (scan margin x staging pool x row
    sort width x dense-scan variant x one or two launches per frame, the
    spring-clumps winner among them) runs the
    oracle's frames bit for bit on a uniform scene, a scene with dense rows
    (bucket spill, rows sorted in global memory, unstaged collide blocks) and
    a spring-clumps scene."""
    from scene_helpers import dense_row_scene

    if scene == "uniform":
        cfg = replace(CONFIGS["uniform_1m"], number_of_agents=20_000, grids=128)
        ps = cfg.settings()
        pos, vf = generate(cfg, seed=9)
    elif scene == "clumps":
        cfg = replace(CONFIGS["spring_clumps_1m"], number_of_agents=65_536, grids=256, clumps=8)
        ps = cfg.settings()
        pos, vf = generate(cfg, seed=9)
    else:
        ps, pos, vf = dense_row_scene()
    r = oracle.World(ps)
    r.upload(pos, vf)
    r.step(6)
    pr, vr = r.download()
    kr = r.debug()["next_keys"]
    r.close()
    for tuning in TUNINGS:
        print(f"[tuning] {scene} {tuning}", flush=True)
        with Physics(ps) as g:
            g.set_tuning(*tuning)
            assert g.tuning == tuning
            g.upload_state(pos, vf)
            g.run(6)
            pg, vg = g.download_state()
            assert np.array_equal(pg.view(np.uint32), pr.view(np.uint32)), tuning
            assert np.array_equal(vg, vr), tuning
            assert np.array_equal(g.debug()["next_keys"], kr), tuning


@pytest.mark.parametrize("spring", [False, True])
@pytest.mark.parametrize("tuning", [(1.0, 896, 512, 0), (1.0, 1408, 512, 0),
                                    (0.0625, 1408, 512, 1)])
def test_window_staging_batches(gpu, oracle, spring, tuning):
    """A collide block stages its windows four slots per thread at a time
    (1024 slots a batch).  Clumps of 32,768 spheres on a 64 x 64 grid give
    blocks whose windows hold from ~800 to ~1400 spheres: with the 896-slot
    pool some overflow it and read from global memory; with the 1408-slot
    pool some need a second, partly used batch.  Rigid and spring responses
    (the rigid blocks also stage each slot's input index)."""
    cfg = replace(CONFIGS["spring_clumps_1m"], number_of_agents=32_768, grids=64, clumps=4,
                  enable_spring=spring)
    print(f"[staging] spring={spring} {tuning}", flush=True)
    g, r = make_pair(oracle, cfg, seed=3)
    with g:
        g.set_tuning(*tuning)
        compare_frames(g, r, 12, check_every=4)
    r.close()


@pytest.mark.parametrize("dense", [0, 1])
@pytest.mark.parametrize("pool", [896, 1408])
def test_large_scene_collide_variant_bit_exact(gpu, oracle, pool, dense):
    """Scenes above 2^21 spheres launch the collide variant that raises a
    block's wave priority while it stages its windows; its spring and plain
    instantiations, with either pool, must give the oracle's frames."""
    n = (1 << 21) + 4096
    for name in ("spring_clumps_1m", "uniform_1m"):
        cfg = replace(CONFIGS[name], number_of_agents=n, grids=1448)
        print(f"[large] {name} n={n} pool={pool}", flush=True)
        g, r = make_pair(oracle, cfg, seed=13)
        with g:
            g.set_tuning(1.0, pool, 512, dense)
            compare_frames(g, r, 3)
        r.close()


def test_autotune_restores_state_and_keeps_results(gpu):
    cfg = replace(CONFIGS["uniform_1m"], number_of_agents=20_000, grids=128)
    ps = cfg.settings()
    pos, vf = generate(cfg, seed=3)
    with Physics(ps) as a, Physics(ps) as b:
        for g in (a, b):
            g.upload_state(pos, vf)
            g.run(3)
        before = a.download_state()
        best, ms = a.autotune(2)
        assert best in TUNINGS and ms > 0
        assert a.tuning == best and a.time() == b.time() == 2
        after = a.download_state()
        assert np.array_equal(before[0].view(np.uint32), after[0].view(np.uint32))
        assert np.array_equal(before[1], after[1])
        a.run(4)
        b.run(4)
        pa, va = a.download_state()
        pb, vb = b.download_state()
        assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)) and np.array_equal(va, vb)
    with pytest.raises(Exception, match="scan_margin"):
        with Physics(ps) as c:
            c.set_tuning(0.0, 896)
    with pytest.raises(Exception, match="stage_pool"):
        with Physics(ps) as c:
            c.set_tuning(1.0, 1000)
    with pytest.raises(Exception, match="row_sort_threads"):
        with Physics(ps) as c:
            c.set_tuning(1.0, 896, 128)


def test_checkpoint_rejects_other_sizes(gpu):
    cfg = replace(CONFIGS["plain_65k"], number_of_agents=4096, grids=64)
    with Physics(cfg.settings()) as a:
        a.upload_state(*generate(cfg, seed=1))
        ck = a.save_checkpoint()
    other = replace(cfg, number_of_agents=8192)
    with Physics(other.settings()) as b:
        with pytest.raises(Exception, match="4096"):
            b.load_checkpoint(ck)


@pytest.mark.parametrize("grids,grid_size,n,mass_mode", [
    (64, 3, 4096, "ones"),        # divisor 3: the general fast-division path
    (100, 1, 10_000, "ones"),     # G not a power of two: row = key / 100
    (64, 1, 4096, "random"),      # masses off powers of two: IEEE divisions
    (48, 2, 2304, "pow2"),        # masses 0.5 / 2 / 4: exact reciprocals
])
def test_non_power_of_two_paths(gpu, oracle, grids, grid_size, n, mass_mode):
    rng = np.random.default_rng(grids * 7 + grid_size)
    ps = PhysicsSettings(number_of_grids_one_dimension=grids, grid_size=grid_size,
                         half_map_height=6, e_agent=0.7, e_envir=0.6, time_delta=1 / 60,
                         _max_number_of_agents=n)
    hms = (grids * grid_size) // 2
    pos = np.empty((n, 4), np.float32)
    pos[:, 0] = rng.uniform(-hms + 0.5, hms - 0.5, n)
    pos[:, 2] = rng.uniform(-hms + 0.5, hms - 0.5, n)
    pos[:, 1] = rng.uniform(0.5, 11.5, n)
    pos[:, 3] = rng.uniform(0.3, 0.7, n) if mass_mode == "random" else 0.5
    if mass_mode == "random":
        mass = rng.uniform(0.3, 3.0, n).astype(np.float32)
    elif mass_mode == "pow2":
        mass = rng.choice(np.float32([0.5, 1.0, 2.0, 4.0]), n)
    else:
        mass = np.ones(n, np.float32)
    force = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    vf = pack_vfme(rng.uniform(-1, 1, (n, 3)).astype(np.float32), force, mass)
    g = Physics(ps)
    g.upload_state(pos, vf)
    r = oracle.World(ps)
    r.upload(pos, vf)
    with g:
        compare_frames(g, r, 12)
