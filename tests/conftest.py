import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); parity tests proper")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def hip_ctx():
    """One device context for the GPU tests (fails loudly without the library)."""
    from disco_amd import backend as B
    if not B.gpu_available():
        pytest.fail("no usable HIP device")
    ctx = B.Context(0)
    yield ctx
    ctx.close()
