"""Scenes shared by the GPU tests."""
import numpy as np

from bevy_millions_ball_amd import PhysicsSettings
from bevy_millions_ball_amd.scenes import pack_vfme

# rows far above the mean row (256 spheres): every one spills past the bucket
# capacity (1024); row 5 also exceeds the 4096 spheres a row sort keeps in LDS;
# rows 20-22 make the collide windows of their blocks overflow the LDS pool
DENSE_ROWS = ((5, 5000), (20, 3000), (21, 3000), (22, 3000))


def dense_row_scene(n=16384, grids=64, dense=DENSE_ROWS, seed=21):
    rng = np.random.default_rng(seed)
    ps = PhysicsSettings(number_of_grids_one_dimension=grids, grid_size=1, half_map_height=8,
                         e_agent=0.8, e_envir=0.8, time_delta=1 / 60, _max_number_of_agents=n)
    hms = grids // 2
    pos = np.empty((n, 4), np.float32)
    i = 0
    for row, count in dense:
        z0 = -hms + row
        pos[i:i + count, 2] = rng.uniform(z0 + 0.01, z0 + 0.99, count)
        pos[i:i + count, 0] = rng.uniform(-hms + 0.5, hms - 0.5, count)
        i += count
    pos[i:, 0] = rng.uniform(-hms + 0.5, hms - 0.5, n - i)
    pos[i:, 2] = rng.uniform(-hms + 0.5, hms - 0.5, n - i)
    pos[:, 1] = rng.uniform(0.5, 15.5, n)
    pos[:, 3] = 0.5
    pos = pos[rng.permutation(n)]
    vf = pack_vfme(rng.uniform(-1, 1, (n, 3)).astype(np.float32), np.zeros((n, 3), np.float32),
                   np.ones(n, np.float32))
    return ps, pos, vf
