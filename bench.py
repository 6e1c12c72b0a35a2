"""Throughput of the physics step on one MI355X.

Workload (default): 1,048,576 spheres, 1024 x 1024 cells of size 1, half map
height 8, radius 0.5, dt 1/60, e_agent = e_envir = 0.8, spring off, seeded
uniform positions inside the box and velocities U[-1,1]^3 (per-step units).
A "step" is one LOOP frame (stable sort of the cell keys with the per-cell
ranges, then collide + integrate) of that scene with its state resident in
HBM; the INIT frame and the warmup frames run before the timed region.

Prints one JSON line (rank 0).  `roofline` is computed for the collide +
integrate kernel from its HIP-event durations over a profiled run of the same
K frames; `cpu_baseline` times the oracle (the C transcription, OpenMP) on a
bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from dataclasses import replace
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "sphere-steps/sec at 1M spheres; p50/p99 step latency, mixed-scene load"
HBM_PEAK_GBS = 8000.0


def physics_bytes_per_sphere(n: int, cells: int) -> float:
    """Algorithmic HBM bytes of one collide + integrate launch per sphere:
    sorted key in 4 + perm 4 + state in 32 + cell offsets 4 * C / N
    + state out 32 + next key out 4 (written over the sorted key)
    + (column, slot) pair filed in the next row's bucket 8."""
    return 84.0 + 4.0 * cells / n


def row_sort_bytes_per_sphere(n: int, cells: int) -> float:
    """Row sort: bucket pair in 8 + perm out 4 + sorted key out 4 + cell
    offsets out 4 * C / N."""
    return 16.0 + 4.0 * cells / n


def read_traffic(name: str):
    """HBM bytes per collide + integrate launch from the counter summary
    profiles/traffic_<workload>.txt (`key = value` lines), or None."""
    path = ROOT / "profiles" / f"traffic_{name}.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        key, sep, val = line.partition("=")
        if sep and key.strip() == "physics_step_kernel_bytes_per_launch":
            return float(val.strip())
    return None


def cpu_frames(cfg, pos, vf, budget_s: float, threads: int):
    """LOOP frames of the oracle after INIT for about budget_s seconds:
    (frames, seconds, per-frame seconds of [sort, spatial hashing, physics])."""
    from oracle import mb_oracle

    w = mb_oracle.World(cfg.settings())
    w.upload(pos, vf)
    w.step(1, threads=threads)  # INIT frame, not counted
    stages = [0.0, 0.0, 0.0]
    frames = 0
    t0 = time.perf_counter()
    while True:
        w.step(1, threads=threads, stage_seconds=stages)
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s or frames >= 2000:
            break
    w.close()
    return frames, el, [s / frames for s in stages]


def cpu_baseline(cfg, pos, vf, budget_s: float, threads: int) -> dict:
    """The oracle (the C transcription of the reference's frame) on the host:
    OpenMP over `threads` cores for budget_s seconds, then one thread for
    half that, each with its stage split."""
    frames, el, stages = cpu_frames(cfg, pos, vf, budget_s, threads)
    f1, el1, stages1 = cpu_frames(cfg, pos, vf, budget_s / 2.0, 1)
    names = ("sort", "spatial_hashing", "physics")
    n = cfg.number_of_agents
    return {
        "value": n * frames / el,
        "unit": "sphere-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{cfg.name}: {n} spheres, {frames} LOOP frames after INIT, "
                  f"{el:.1f} s, C transcription with OpenMP",
        "stage_ms_per_frame": {k: v * 1e3 for k, v in zip(names, stages)},
        "single_thread": {
            "value": n * f1 / el1,
            "sample": f"{f1} LOOP frames after INIT, {el1:.1f} s, one thread",
            "stage_ms_per_frame": {k: v * 1e3 for k, v in zip(names, stages1)},
        },
    }


def serving_bench(n_requests: int, n_poisson: int, seed: int) -> dict:
    """Batched serving of the mixed-scene workload against stepping the same
    requests one by one on the same GPU, then latency under Poisson load.

    Requests: N log-uniform in [4096, 1,048,576], G = next pow2 of ceil(sqrt N),
    1-60 LOOP steps, plain-step distributions (bevy_millions_ball_amd.serving).
    The timed region of both throughput legs is the server's work after the
    requests were handed over: admission, host->device upload, INIT + LOOP
    frames, device->host retirement.  Throughput counts LOOP sphere-steps.
    The batched leg admits first come, first served; the most-steps-first
    order is reported beside it."""
    from bevy_millions_ball_amd.serving import Server, poisson_requests, request_scene

    reqs = poisson_requests(n_requests, rate_per_s=200.0, seed=seed)
    scenes = [request_scene(r) for r in reqs]
    work = float(sum(r.number_of_agents * r.steps for r in reqs))

    # warm up the kernels and the pinned pool on one small scene
    with Server() as srv:
        st, pos, vf = scenes[0]
        rid = srv.submit(st, pos, vf, 2)
        srv.drain()
        srv.collect(rid)

    def batched(admission):
        with Server(admission=admission) as srv:
            ids = [srv.submit(st, pos, vf, r.steps) for (st, pos, vf), r in zip(scenes, reqs)]
            t0 = time.perf_counter()
            srv.drain()
            el = time.perf_counter() - t0
            frames = srv.stats()["frames"]
            for i in ids:
                srv.collect(i)
        return el, frames

    batched_s, frames_batched = batched("fifo")
    lpt_s, frames_lpt = batched("most_steps_first")
    # the same batch again with events around every frame, for device time
    with Server(profile_frames=True) as srv:
        ids = [srv.submit(st, pos, vf, r.steps) for (st, pos, vf), r in zip(scenes, reqs)]
        srv.drain()
        batched_device_s = float(srv.frame_times().sum()) * 1e-3
        for i in ids:
            srv.collect(i)

    one_s = 0.0
    with Server() as srv:
        for (st, pos, vf), r in zip(scenes, reqs):
            rid = srv.submit(st, pos, vf, r.steps)
            t0 = time.perf_counter()
            srv.drain()
            one_s += time.perf_counter() - t0
            srv.collect(rid)
        frames_one = srv.stats()["frames"]

    # Poisson arrivals at 200/s: submit at arrival time, tick while work remains
    preqs = poisson_requests(n_poisson, rate_per_s=200.0, seed=seed + 1)
    pscenes = [request_scene(r) for r in preqs]
    lat = []
    with Server(profile_frames=True) as srv:
        t_start = time.perf_counter()
        nxt = 0
        live = []
        while nxt < len(preqs) or live:
            now = time.perf_counter() - t_start
            while nxt < len(preqs) and preqs[nxt].arrival_s <= now:
                st, pos, vf = pscenes[nxt]
                live.append(srv.submit(st, pos, vf, preqs[nxt].steps))
                nxt += 1
            if srv.tick() == 0 and nxt < len(preqs):
                time.sleep(max(0.0, min(0.0005, preqs[nxt].arrival_s - now)))
            for rid in list(live):
                stt = srv.status(rid)
                if stt["state"] == "done":
                    lat.append((stt["done_s"] - stt["submit_s"]) * 1e3)
                    srv.collect(rid)
                    live.remove(rid)
        wall = time.perf_counter() - t_start
        frame_ms = srv.frame_times()
    pwork = float(sum(r.number_of_agents * r.steps for r in preqs))
    lat = np.array(lat)
    return {
        "workload": "mixed scenes: N log-uniform [4096, 1048576], G = next_pow2(ceil(sqrt N)), "
                    "1-60 LOOP steps, plain-step distributions",
        "requests": n_requests,
        "spheres_total": int(sum(r.number_of_agents for r in reqs)),
        "sphere_steps": work,
        "admission": "fifo",
        "batched_sphere_steps_per_s": work / batched_s,
        "one_by_one_sphere_steps_per_s": work / one_s,
        "batched_over_one_by_one": one_s / batched_s,
        "batched_s": batched_s,
        "batched_frames_device_s": batched_device_s,
        "one_by_one_s": one_s,
        "frames_batched": int(frames_batched),
        "most_steps_first": {"batched_s": lpt_s, "frames": int(frames_lpt),
                             "batched_over_one_by_one": one_s / lpt_s},
        "frames_one_by_one": int(frames_one),
        "poisson": {
            "rate_per_s": 200.0,
            "requests": n_poisson,
            "wall_s": wall,
            "sphere_steps_per_s": pwork / wall,
            "request_latency_ms_p50": float(np.percentile(lat, 50)),
            "request_latency_ms_p99": float(np.percentile(lat, 99)),
            "frame_ms_p50": float(np.percentile(frame_ms, 50)),
            "frame_ms_p99": float(np.percentile(frame_ms, 99)),
            "frames": int(len(frame_ms)),
        },
    }


def serving_trace(n_requests: int, seed: int) -> int:
    """One batched drain (first come, first served) of the serving mix, after
    a warm-up server; prints its wall time.  Made to run under a kernel trace."""
    import torch

    from bevy_millions_ball_amd.serving import Server, poisson_requests, request_scene

    torch.cuda.set_device(0)
    reqs = poisson_requests(n_requests, rate_per_s=200.0, seed=seed)
    scenes = [request_scene(r) for r in reqs]
    with Server() as srv:
        st, pos, vf = scenes[0]
        rid = srv.submit(st, pos, vf, 2)
        srv.drain()
        srv.collect(rid)
    with Server() as srv:
        ids = [srv.submit(st, pos, vf, r.steps) for (st, pos, vf), r in zip(scenes, reqs)]
        t0 = time.perf_counter()
        srv.drain()
        el = time.perf_counter() - t0
        frames = srv.stats()["frames"]
        for i in ids:
            srv.collect(i)
    print(json.dumps({"serving_trace": {"requests": n_requests, "batched_s": el, "frames": int(frames)}}),
          flush=True)
    return 0


def copy_calibration(launches: int = 5) -> int:
    """A plain device copy of known size (64 MiB read + 64 MiB written per
    launch), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: known bytes
    divided by the counter's KiB scale the kernels' counters to bytes."""
    import torch

    copy_bytes = 64 * 1024 * 1024
    torch.cuda.set_device(0)
    x = torch.empty(copy_bytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(launches):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"launches": launches, "bytes_read": copy_bytes,
                                           "bytes_written": copy_bytes}}), flush=True)
    return 0


def other_config(name: str, seed: int, steps: int, cpu_seconds: float = 0.0,
                 cpu_threads: int = 1) -> dict:
    """Throughput and per-kernel device time of another workload (the
    parity-test configurations), `steps` LOOP frames after INIT + 10 and the
    autotuner's trials; with cpu_seconds > 0 also the oracle on the host for
    a sample of about that long (and a one-thread sample half as long)."""
    import torch

    from bevy_millions_ball_amd import CONFIGS, Physics, generate

    cfg = CONFIGS[name]
    n = cfg.number_of_agents
    cells = cfg.grids * cfg.grids
    pos, vf = generate(cfg, seed=seed)
    with Physics(cfg.settings()) as phys:
        phys.upload_state(pos, vf)
        phys.run(11)
        # the library's autotuner picks the launch choices (never the results)
        tuning, _ = phys.autotune(frames_per_trial=5)
        graph = phys.frame_graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        phys.run(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        phys.set_profiling(True)
        phys.run(steps)
        torch.cuda.synchronize()
        stage_ms = phys.profile_read().mean(axis=0)
        phys.set_profiling(False)
    achieved = physics_bytes_per_sphere(n, cells) * n / (stage_ms[1] * 1e-3) / 1e9
    cpu = cpu_baseline(cfg, pos, vf, cpu_seconds, cpu_threads) if cpu_seconds > 0 else None
    return {
        "spheres": n, "grids": cfg.grids, "spring": cfg.enable_spring, "layout": cfg.layout,
        "time_delta": cfg.time_delta, "steps": steps,
        "tuning": {"scan_margin": tuning[0], "stage_pool": tuning[1],
                   "row_sort_threads": tuning[2], "dense_scan": tuning[3],
                   "one_launch": tuning[4], "frame_graph": graph},
        "value": n * steps / el, "unit": "sphere-steps/s", "ms_per_step": el * 1e3 / steps,
        "stage_ms": {"row_sort": float(stage_ms[0]), "collide_integrate": float(stage_ms[1])},
        "collide_roofline_frac": achieved / HBM_PEAK_GBS,
        "cpu_baseline": cpu,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="uniform_1m")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--latency-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=64)
    ap.add_argument("--poisson-requests", type=int, default=200)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true")
    ap.add_argument("--serving-trace", action="store_true",
                    help="only one batched drain of the serving mix (for a kernel trace)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the library's autotuner pick the launch choices before timing")
    ap.add_argument("--tuning", default=None,
                    help="scan_margin,stage_pool,row_sort_threads[,dense_scan[,one_launch]] "
                         "for the measured workload")
    ap.add_argument("--frame-graph", action="store_true",
                    help="replay runs of frames as a captured graph instead of launching each kernel")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a plain 64 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent scenes shard by scene, one process per GPU).", file=sys.stderr)
        return 2

    import torch

    from bevy_millions_ball_amd import CONFIGS, Physics, generate

    if args.serving_trace:
        return serving_trace(args.serving_requests, args.seed)
    if args.copy_calibration:
        return copy_calibration()

    cfg = CONFIGS[args.config]
    n = cfg.number_of_agents
    cells = cfg.grids * cfg.grids
    pos, vf = generate(cfg, seed=args.seed)
    torch.cuda.set_device(0)
    phys = Physics(cfg.settings())
    if args.frame_graph:
        phys.set_frame_graph(True)
    if args.tuning:
        m, p, s, d, o = (args.tuning.split(",") + ["0", "0"])[:5]
        phys.set_tuning(float(m), int(p), int(s), int(d), int(o))
    t0 = time.perf_counter()
    phys.upload_state(pos, vf)  # host -> HBM, synchronised by the library
    upload_ms = (time.perf_counter() - t0) * 1e3
    phys.run(1)  # INIT frame
    phys.run(args.warmup)
    tuning = None
    if args.autotune:
        # the library times its candidate launch choices on a copy of the
        # state and keeps the fastest (results are the same for all of them)
        tuning, _ = phys.autotune(frames_per_trial=10)
        tuning = tuning + (phys.frame_graph,)
    torch.cuda.synchronize()

    # -- throughput: K frames back to back -----------------------------------
    t0 = time.perf_counter()
    phys.run(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = n * args.steps / elapsed

    # -- per-kernel-family device time over K profiled frames ---------------
    phys.set_profiling(True)
    phys.run(args.steps)
    torch.cuda.synchronize()
    prof = phys.profile_read()
    phys.set_profiling(False)
    stage_ms = prof.mean(axis=0)
    device_step_ms = prof.sum(axis=1)

    # -- latency: each frame launched and waited for on its own ---------------
    lat = []
    for _ in range(args.latency_steps):
        t = time.perf_counter()
        phys.run(1)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat)
    t0 = time.perf_counter()
    phys.download_state()  # HBM -> host, synchronised by the library
    download_ms = (time.perf_counter() - t0) * 1e3
    phys.close()

    phys_bytes = physics_bytes_per_sphere(n, cells) * n
    achieved = phys_bytes / (stage_ms[1] * 1e-3) / 1e9
    rsort_gbs = row_sort_bytes_per_sphere(n, cells) * n / (stage_ms[0] * 1e-3) / 1e9

    traffic = read_traffic(cfg.name)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "sphere-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": cfg.name,
            "spheres": n,
            "grids": cfg.grids,
            "grid_size": cfg.grid_size,
            "half_map_height": cfg.half_map_height,
            "radius": cfg.radius,
            "time_delta": cfg.time_delta,
            "e_agent": cfg.e_agent,
            "e_envir": cfg.e_envir,
            "spring": cfg.enable_spring,
            "layout": cfg.layout,
            "seed": args.seed,
            "parallelism": "single GPU",
            "tuning": None if tuning is None else {
                "scan_margin": tuning[0], "stage_pool": tuning[1], "row_sort_threads": tuning[2],
                "dense_scan": tuning[3], "one_launch": tuning[4], "frame_graph": tuning[5]},
        },
        "p50_step_ms": float(np.percentile(lat, 50)),
        "p99_step_ms": float(np.percentile(lat, 99)),
        "upload_ms": upload_ms,
        "download_ms": download_ms,
        "device_step_ms_p50": float(np.percentile(device_step_ms, 50)),
        "device_step_ms_p99": float(np.percentile(device_step_ms, 99)),
        "stage_ms": {"row_sort": float(stage_ms[0]), "collide_integrate": float(stage_ms[1])},
        "stage_achieved_GBs": {"row_sort": rsort_gbs, "collide_integrate": achieved},
        "roofline": {
            "kernel": "physics_step_kernel (collide + integrate)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": phys_bytes,
        },
    }
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not args.no_other_configs and args.config == "uniform_1m":
        # each other workload's CPU sample: 40 % of the headline's budget
        other_cpu = 0.0 if args.no_cpu_baseline else 0.4 * args.cpu_seconds
        line["other_configs"] = {
            name: other_config(name, args.seed, steps, other_cpu, threads)
            for name, steps in (("plain_65k", 500), ("spring_clumps_1m", 100), ("uniform_4m", 100))}
    if not args.no_serving:
        line["serving"] = serving_bench(args.serving_requests, args.poisson_requests, args.seed)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cfg, pos, vf, args.cpu_seconds, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
