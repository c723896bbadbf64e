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


def same_cfar_events(got, exp, what=""):
    """CFAR event records equal: type, id, flags, span and hit count exactly;
    t0 / t1 exactly (frame times); the centroid within 1e-9 and the dB figures
    within 2e-6 relative (power formed by another exp differs by an ulp)."""
    assert len(got) == len(exp), (what, len(got), len(exp), got[["type", "id", "lo_bin", "hi_bin"]][:20],
                                  exp[["type", "id", "lo_bin", "hi_bin"]][:20])
    for f in ("type", "id", "flags", "lo_bin", "hi_bin", "n_hits"):
        assert np.array_equal(got[f], exp[f]), (what, f, got[f], exp[f])
    for f in ("t0", "t1"):
        assert np.array_equal(got[f], exp[f], equal_nan=True), (what, f, got[f], exp[f])
    np.testing.assert_allclose(got["fc_bin"], exp["fc_bin"], rtol=1e-9, atol=1e-9, err_msg=what)
    for f in ("level_db", "noise_db", "snr_db", "band_snr_db"):
        np.testing.assert_allclose(got[f], exp[f], rtol=2e-6, atol=2e-5, err_msg=f"{what} {f}")
