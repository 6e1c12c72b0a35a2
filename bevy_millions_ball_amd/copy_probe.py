"""Plain device copy of known size, for calibrating the HBM traffic counters.

rocprofv3's FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are read for this
copy (64 MiB read + 64 MiB written per launch) and for the physics kernels
under the same command line options; the ratio known bytes / counter for the
copy scales the kernels' counters to bytes.

    rocprofv3 --pmc FETCH_SIZE -- python3 bevy_millions_ball_amd/copy_probe.py
"""
from __future__ import annotations

import torch

COPY_BYTES = 64 * 1024 * 1024
LAUNCHES = 5


def main() -> None:
    x = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(LAUNCHES):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy probe: {LAUNCHES} launches, {COPY_BYTES} bytes read and {COPY_BYTES} written each",
          flush=True)


if __name__ == "__main__":
    main()
