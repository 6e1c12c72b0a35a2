"""Long bit-exact runs against the oracle, by hand (about 70 s on one MI355X;
not collected by pytest): 1000 LOOP frames of uniform_1m, 300 of
spring_clumps_1m at the point the autotuner picks for it, 200 of uniform_4m,
each compared with the oracle's frames every 50 or 100 frames (state, packed
f16 halves and next keys, bit for bit).

    python tests/long_parity_check.py
"""
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import numpy as np
from bevy_millions_ball_amd import Physics, CONFIGS, generate
from oracle import mb_oracle

def run(name, frames, every, tuning=None, seed=11):
    cfg = CONFIGS[name]
    pos, vf = generate(cfg, seed=seed)
    g = Physics(cfg.settings())
    if tuning: g.set_tuning(*tuning)
    g.upload_state(pos, vf)
    w = mb_oracle.World(cfg.settings()); w.upload(pos, vf)
    t0 = time.time(); checks = 0
    for f in range(0, frames + 1, every):
        n = every if f else 1  # INIT first
        g.run(n); w.step(n)
        pg, vg = g.download_state(); pr, vr = w.download()
        ok = np.array_equal(pg.view(np.uint32), pr.view(np.uint32)) and np.array_equal(vg, vr) \
            and np.array_equal(g.debug()["next_keys"], w.debug()["next_keys"])
        checks += 1
        if not ok:
            print(f"{name}: MISMATCH at frame {g.time()}", flush=True); return False
    print(f"{name}: {frames} LOOP frames, state and keys bit-identical at all {checks} checks "
          f"(every {every} frames), {time.time() - t0:.0f} s", flush=True)
    g.close(); w.close()
    return True

ok = run("uniform_1m", 1000, 100)
ok = run("spring_clumps_1m", 300, 50, tuning=(0.0625, 1408, 512, 1)) and ok
ok = run("uniform_4m", 200, 50) and ok
sys.exit(0 if ok else 1)
