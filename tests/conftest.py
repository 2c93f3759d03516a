import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def known():
    return json.loads((GOLDEN / "known_answers.json").read_text())


@pytest.fixture(scope="session")
def numpy_ref():
    return json.loads((GOLDEN / "numpy_reference.json").read_text())


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def gpu():
    """Loads libnumrs_amd.so and selects device 0. Fails (never skips) if the library is missing."""
    import numrs_amd
    from numrs_amd import device

    device.init(0)
    return numrs_amd


def fval(v):
    """JSON value -> float ('nan', 'inf', '-inf' strings allowed)."""
    return float(v) if not isinstance(v, str) else float(v)


def det_data(dtype, n=50000):
    """tests/test_reduction_determinism.rs:7-14: x_i = i * 0.5 - 12345."""
    i = np.arange(n)
    if dtype == np.float32:
        return (i.astype(np.float32) * np.float32(0.5) - np.float32(12345.0)).astype(np.float32)
    return i.astype(np.float64) * 0.5 - 12345.0
