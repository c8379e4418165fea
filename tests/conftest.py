import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    from oracle import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    return load


@pytest.fixture(scope="session")
def ggrid():
    g = dict(np.load(os.path.join(GOLDEN, "grid.npz")))
    return {k: (int(v) if k in ("nx", "ny", "nz", "ng") else float(v)) for k, v in g.items()}


@pytest.fixture(scope="session", autouse=True)
def _tuning_from_env():
    """PXR_TEST_TUNE="key=value,key=value" runs the tests with those launch
    parameters of the tiled kernels (pxrhip_set_tuning), e.g. to cover a
    non-default form of the fused step."""
    spec = os.environ.get("PXR_TEST_TUNE", "")
    if spec:
        from picsar_amd import _lib
        for kv in spec.split(","):
            k, v = kv.split("=")
            _lib.call("pxrhip_set_tuning", k.strip().encode(), int(v))
    yield
