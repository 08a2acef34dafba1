import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"

# Tower x of the reference's single-tower fixtures: equirectangular projection
# of (lat 50.95, lon 11.5867) about (50.95, 11.586) with R = 6371 km
# (tests/conftest.py:41-67 of the reference; pinned by fitting the fixture).
TOWER_X = float(np.deg2rad(11.5867 - 11.586) * 6371000.0 * np.cos(np.deg2rad(50.95)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


def load_golden(name):
    return dict(np.load(GOLDEN / f"{name}.npz"))


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
