"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def pkg():
    return ge.load_package()


@pytest.fixture(scope="session")
def orc():
    return ge.load_oracle()


def rowwise_rel_err(a, b, floor_frac=1e-2):
    """Per-particle normwise error ||a_i - b_i|| / max(||b_i||, floor), where
    floor = floor_frac x RMS of ||b|| over the field (vectors near zero are
    compared in absolute terms against the field scale)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 1:
        a, b = a[:, None], b[:, None]
    nb = np.linalg.norm(b, axis=1)
    floor = floor_frac * np.sqrt(np.mean(nb ** 2)) if len(nb) else 0.0
    den = np.maximum(nb, max(floor, 1e-30))
    return np.linalg.norm(a - b, axis=1) / den
