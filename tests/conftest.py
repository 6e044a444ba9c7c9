import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the GPU)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.build()
    return o


@pytest.fixture(scope="session")
def gpu():
    """The loaded C-ABI library; skips when no GPU is visible."""
    from fuel_amd import _lib
    lib = _lib.load()
    if _lib.device_count() < 1:
        pytest.skip("no GPU visible")
    return _lib


def load_gaussian4d():
    """testdata/6-gaussian-4d.csv loader (src/outlier/common.rs:200-229)."""
    pts = []
    with open(os.path.join(GOLDEN, "6-gaussian-4d.csv")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            parts = line.split()
            if len(parts) < 5:
                continue
            pts.append([float(x) for x in parts[:4]])
    return np.array(pts, dtype=np.float64)


def load_reference_scores():
    """testdata/reference-outlier-scores.csv loader (src/outlier/common.rs:231-254)."""
    out = {}
    with open(os.path.join(GOLDEN, "reference-outlier-scores.csv")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            parts = line.split()
            out[parts[0]] = np.array([float(x) for x in parts[1:]])
    return out
