import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def gpu_available():
    try:
        import torch  # noqa: F401  (only to ask the runtime whether a device exists)

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_explicit():
    from mpcns_uns_amd import workloads

    return workloads.build(net=(12, 48, 3), hall=False)


@pytest.fixture(scope="session")
def small_hall():
    from mpcns_uns_amd import workloads

    return workloads.build(net=(12, 48, 3), hall=True)


def maxrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300)) if b.size else 0.0
