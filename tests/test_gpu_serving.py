"""GPU tests of the serving path (mb_server_*): many scenes in one frame.

Each scene's results must be the same alone and in any mix, bit for bit
(state, the perm of its last frame and its next keys), and equal to the
single-scene path and to the oracle on the same inputs."""
from dataclasses import replace

import numpy as np
import pytest

from bevy_millions_ball_amd import Physics
from bevy_millions_ball_amd.scenes import SceneConfig, generate
from bevy_millions_ball_amd.serving import Server, scene_grid
from scene_helpers import dense_row_scene

pytestmark = pytest.mark.gpu


def scene(n, steps, seed, grids=None, **kw):
    g = grids if grids is not None else scene_grid(n)
    cfg = SceneConfig(f"t{n}", n, g, **kw)
    pos, vf = generate(cfg, seed=seed)
    return cfg.settings(), pos, vf, steps


# sizes, grids and modes that exercise the batched kernels' scene lookups:
# G^2 > N, G^2 == N, spring on, a single sphere, INIT only, a scene spanning
# several partition tiles and collide blocks with a partial last block, and
# crowded scenes of both modes (the dense collide variant)
MIX = [
    scene(4096, 5, 1),
    scene(5000, 3, 2),                      # G = 128, G^2 > N
    scene(20000, 7, 3),                     # G = 256
    scene(65536, 4, 4, grids=256),          # G^2 == N
    scene(30000, 6, 5, grids=128, enable_spring=True, k_agent=500.0, c_agent=5.0,
          time_delta=1.0 / 120.0, layout="clumps", clumps=4, clump_sigma_cells=6.0),
    scene(16000, 5, 15, grids=128, layout="clumps", clumps=3, clump_sigma_cells=5.0),  # crowded, rigid
    scene(1, 2, 6, grids=2),                # one sphere, G^2 > N
    scene(70001, 0, 7),                     # INIT only, G = 512
    scene(9999, 9, 8, grids=100),           # G not a power of two
    dense_row_scene() + (3,),               # bucket spill, global row sort, unstaged collide
]


def run_alone(items, **server_kw):
    out = []
    with Server(**server_kw) as srv:
        for st, pos, vf, steps in items:
            rid = srv.submit(st, pos, vf, steps, keep_debug=True)
            srv.drain()
            out.append(srv.collect(rid))
    return out


def run_mix(items, **server_kw):
    with Server(**server_kw) as srv:
        ids = [srv.submit(st, pos, vf, steps, keep_debug=True) for st, pos, vf, steps in items]
        srv.drain()
        stats = srv.stats()
        return [srv.collect(i) for i in ids], stats


def assert_same(a, b, label):
    assert np.array_equal(a.position_with_scale.view(np.uint32),
                          b.position_with_scale.view(np.uint32)), f"{label}: positions"
    assert np.array_equal(a.packed_vfme, b.packed_vfme), f"{label}: vfme"
    assert np.array_equal(a.perm, b.perm), f"{label}: perm"
    assert np.array_equal(a.next_keys, b.next_keys), f"{label}: next keys"


@pytest.mark.parametrize("use_graph", [False, True])
def test_alone_equals_mix(gpu, use_graph):
    alone = run_alone(MIX, use_graph=use_graph)
    mixed, stats = run_mix(MIX, use_graph=use_graph)
    for i, (a, b) in enumerate(zip(alone, mixed)):
        assert_same(a, b, f"scene {i} (N={MIX[i][0].max_number_of_agents()})")
    assert stats["sphere_steps"] == sum(st.max_number_of_agents() * steps
                                        for st, _, _, steps in MIX)


# a scene of ~2200 spheres per row: the batched row sort takes its wide blocks
DENSE_MIX = [
    scene(70000, 3, 11, grids=32),
    scene(4096, 4, 12),
    scene(20000, 2, 13, enable_spring=True, k_agent=500.0, c_agent=5.0, time_delta=1.0 / 120.0),
]


def test_dense_rows_mix(gpu, oracle):
    alone = run_alone(DENSE_MIX)
    mixed, _ = run_mix(DENSE_MIX)
    for i, (a, b) in enumerate(zip(alone, mixed)):
        assert_same(a, b, f"scene {i}")
    for (st, pos, vf, steps), got in zip(DENSE_MIX, mixed):
        w = oracle.World(st)
        w.upload(pos, vf)
        w.step(1 + steps)
        p2, v2 = w.download()
        w.close()
        assert np.array_equal(p2.view(np.uint32), got.position_with_scale.view(np.uint32))
        assert np.array_equal(v2, got.packed_vfme)


def test_most_steps_first_admission_keeps_results(gpu):
    # admission order changes when a scene starts, never what it computes;
    # a cap that holds few scenes at a time makes the order matter
    alone = run_alone(MIX)
    cap = 3 * max(Server.scene_bytes(st.max_number_of_agents(), st.number_of_grids_one_dimension)
                  for st, _, _, _ in MIX)
    mixed, stats = run_mix(MIX, admission="most_steps_first", memory_cap_bytes=cap)
    for i, (a, b) in enumerate(zip(alone, mixed)):
        assert_same(a, b, f"scene {i} (N={MIX[i][0].max_number_of_agents()})")
    assert stats["sphere_steps"] == sum(st.max_number_of_agents() * steps
                                        for st, _, _, steps in MIX)


def test_admission_order_is_validated(gpu):
    with pytest.raises(ValueError):
        Server(admission="shortest")
    from bevy_millions_ball_amd import _lib
    import ctypes as C
    cfg = _lib.MbServerConfig()
    cfg.memory_cap_bytes = 1 << 30
    cfg.admission = 7
    h = C.c_void_p()
    rc = _lib.lib().mb_server_create(C.byref(cfg), C.byref(h))
    assert rc == _lib.MB_ERR_INVALID_ARGUMENT and not h.value


def test_mix_matches_single_scene_physics(gpu):
    mixed, _ = run_mix(MIX)
    checked = 0
    for (st, pos, vf, steps), got in zip(MIX, mixed):
        n, g = st.max_number_of_agents(), st.number_of_grids_one_dimension
        if g * g > n:
            continue  # the single-scene surface keeps the reference's G^2 <= N rule
        with Physics(st) as ph:
            ph.upload_state(pos, vf)
            ph.run(1 + steps)
            p2, v2 = ph.download_state()
            dbg = ph.debug()
        assert np.array_equal(p2.view(np.uint32), got.position_with_scale.view(np.uint32))
        assert np.array_equal(v2, got.packed_vfme)
        assert np.array_equal(dbg["perm"], got.perm)
        assert np.array_equal(dbg["next_keys"], got.next_keys)
        checked += 1
    assert checked >= 3


def test_mix_matches_oracle(gpu, oracle):
    mixed, _ = run_mix(MIX)
    for (st, pos, vf, steps), got in zip(MIX, mixed):
        w = oracle.World(st)
        w.upload(pos, vf)
        w.step(1 + steps)
        p2, v2 = w.download()
        dbg = w.debug()
        w.close()
        label = f"N={st.max_number_of_agents()} G={st.number_of_grids_one_dimension}"
        assert np.array_equal(p2.view(np.uint32), got.position_with_scale.view(np.uint32)), label
        assert np.array_equal(v2, got.packed_vfme), label
        assert np.array_equal(dbg["perm"], got.perm), label
        assert np.array_equal(dbg["next_keys"], got.next_keys), label


def test_continuous_admission_under_memory_cap(gpu):
    """Requests arrive while others run and wait for memory; results are
    still those of each scene alone."""
    items = [scene(n, s, 100 + i) for i, (n, s) in
             enumerate([(50000, 12), (8000, 3), (120000, 5), (4096, 20), (33000, 8),
                        (90000, 2), (6000, 15), (47000, 6)])]
    alone = run_alone(items)
    sizes = [Server.scene_bytes(st.max_number_of_agents(), st.number_of_grids_one_dimension)
             for st, _, _, _ in items]
    cap = max(sizes) + sorted(sizes)[-2]  # the two largest fit together, not three
    got = {}
    max_queued = 0
    with Server(memory_cap_bytes=cap) as srv:
        ids = []
        pending = list(enumerate(items))
        # three arrive at once, the rest one per tick
        for _ in range(3):
            i, (st, pos, vf, steps) = pending.pop(0)
            ids.append((i, srv.submit(st, pos, vf, steps, keep_debug=True)))
        ticks = 0
        while True:
            if pending:
                i, (st, pos, vf, steps) = pending.pop(0)
                ids.append((i, srv.submit(st, pos, vf, steps, keep_debug=True)))
            srv.tick()
            ticks += 1
            s = srv.stats()
            max_queued = max(max_queued, s["queued"])
            assert s["used_bytes"] <= cap
            for i, rid in list(ids):
                if srv.status(rid)["state"] in ("retiring", "done"):
                    got[i] = srv.collect(rid)
                    ids.remove((i, rid))
            if not pending and not ids:
                break
            assert ticks < 10000
    assert max_queued > 0, "the cap never made a request wait"
    for i, a in enumerate(alone):
        assert_same(a, got[i], f"request {i}")


def test_frame_times_and_status(gpu):
    st, pos, vf, steps = scene(4096, 3, 11)
    with Server(profile_frames=True) as srv:
        rid = srv.submit(st, pos, vf, steps)
        assert srv.status(rid)["state"] == "queued"
        srv.drain()
        status = srv.status(rid)
        assert status["state"] == "done" and status["frames_done"] == 1 + steps
        assert status["submit_s"] <= status["admit_s"] <= status["done_s"]
        times = srv.frame_times()
        assert len(times) == 1 + steps and np.all(times > 0)
        with pytest.raises(Exception):
            srv.collect(rid + 99)
        res = srv.collect(rid)
        assert res.perm is None


def test_submit_rejects_states_outside_the_grid(gpu):
    """The batched frame runs the frame path only: a scene with a sphere
    outside its grid, or radii that let a wall-clamped position leave it, is
    refused at submit (the single-scene handle runs such states through the
    stage kernels instead)."""
    st, pos, vf, _ = scene(4096, 2, 12)
    out = pos.copy()
    out[7, 0] = 40.0                       # map is +-32: column 72
    tiny = pos.copy()
    tiny[:, 3] = 1e-7                      # the wall clamp lands on the +32 edge
    with Server() as srv:
        with pytest.raises(Exception, match="outside the grid"):
            srv.submit(st, out, vf, 2)
        with pytest.raises(Exception, match="leave the grid"):
            srv.submit(st, tiny, vf, 2)
        rid = srv.submit(st, pos, vf, 2)
        srv.drain()
        srv.collect(rid)


def test_submit_reports_the_first_offending_sphere(gpu):
    """A large scene is checked by several host threads, each over a range of
    spheres; whichever thread finds a problem, the message names the first
    offending index in the order of one pass: non-finite values before cells
    outside the grid, and the smallest index of each kind."""
    n = 300_000                                  # four parts of 75,000 spheres
    st, pos, vf, _ = scene(n, 2, 13)
    two_out = pos.copy()
    two_out[250_000, 0] = 1e6                    # fourth part
    two_out[120_000, 2] = 1e6                    # second part: the one to name
    nan_late = two_out.copy()
    nan_late[260_000, 1] = np.nan                # a non-finite float outranks both
    bad_half = vf.copy()
    bad_half[200_000, 1] = 0x7C00                # +inf in an f16 half, third part
    with Server() as srv:
        with pytest.raises(Exception, match="sphere 120000 lies outside"):
            srv.submit(st, two_out, vf, 2)
        with pytest.raises(Exception, match=f"position_with_scale at float {260_000 * 4 + 1}"):
            srv.submit(st, nan_late, vf, 2)
        with pytest.raises(Exception, match=f"packed_vfme at word {200_000 * 4 + 1}"):
            srv.submit(st, pos, bad_half, 2)
        rid = srv.submit(st, pos, vf, 1)
        srv.drain()
        srv.collect(rid)


def test_bench_request_mix_subset(gpu, oracle):
    """The first 10 requests of bench.py's own serving batch (poisson_requests
    with its seed 0: N log-uniform in [4096, 1,048,576], G = next pow2 of
    ceil(sqrt N), 1-60 LOOP steps): each request's result in the mix equals
    its result alone, bit for bit, and the oracle's."""
    from bevy_millions_ball_amd.serving import poisson_requests, request_scene

    reqs = poisson_requests(64, rate_per_s=200.0, seed=0)[:10]
    items = []
    for r in reqs:
        st, pos, vf = request_scene(r)
        items.append((st, pos, vf, r.steps))
    print("[bench mix] N =", [r.number_of_agents for r in reqs], "steps =", [r.steps for r in reqs],
          flush=True)
    alone = run_alone(items)
    mixed, stats = run_mix(items)
    for i, (a, b) in enumerate(zip(alone, mixed)):
        assert_same(a, b, f"request {i} (N={reqs[i].number_of_agents})")
    assert stats["sphere_steps"] == sum(r.number_of_agents * r.steps for r in reqs)
    for (st, pos, vf, steps), got in zip(items, mixed):
        w = oracle.World(st)
        w.upload(pos, vf)
        w.step(1 + steps)
        p2, v2 = w.download()
        assert np.array_equal(w.debug()["next_keys"], got.next_keys)
        w.close()
        assert np.array_equal(p2.view(np.uint32), got.position_with_scale.view(np.uint32))
        assert np.array_equal(v2, got.packed_vfme)
