import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: a long GPU case")


def has_gpu() -> bool:
    """True when a GPU is present; counts devices without initialising HIP
    in this process."""
    return Path("/dev/kfd").exists() and any(Path("/sys/class/kfd/kfd/topology/nodes").glob("*/gpu_id")) \
        and any((p / "gpu_id").read_text().strip() not in ("", "0")
                for p in Path("/sys/class/kfd/kfd/topology/nodes").iterdir() if (p / "gpu_id").exists())


def golden(name: str):
    return np.load(GOLDEN / f"{name}.npz", allow_pickle=False)


@pytest.fixture(scope="session")
def G():
    return golden


@pytest.fixture(scope="session")
def gpu():
    """The product binding, with the library loaded; fails loudly if it is not built."""
    import atkdsp_amd
    atkdsp_amd.load()
    return atkdsp_amd
