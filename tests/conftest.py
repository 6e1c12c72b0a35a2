import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: large CPU case (minutes)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import mb_oracle

    mb_oracle.lib()
    return mb_oracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP device through torch (device memory and streams only)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("GPU test run without a visible GPU")
    from bevy_millions_ball_amd import _lib

    _lib.lib()  # fail loudly if the native library is missing
    return torch.device("cuda:0")
