import json
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "culamp.jl_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libculamp on the GPU)")


# Every GPU test runs under a time limit of its own (pytest-timeout, thread
# method, so that a test stuck inside a GPU call is still ended and named).
# The whole GPU suite takes about 150 s; the limit leaves room for a first
# `import torch` on a fresh machine, which can take a minute or two.
GPU_TEST_SECONDS = 300


def pytest_collection_modifyitems(config, items):
    try:
        import pytest_timeout  # noqa: F401
    except ImportError:
        return
    for item in items:
        if item.get_closest_marker("gpu") is not None and item.get_closest_marker("timeout") is None:
            item.add_marker(pytest.mark.timeout(GPU_TEST_SECONDS, method="thread"))


def unhex(a, shape=None):
    v = np.array([float.fromhex(s) for s in a], dtype=np.float64)
    return v.reshape(shape) if shape is not None else v


def pval(p):
    return math.inf if p == "inf" else float(p)


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "kernels.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="session")
def oracle():
    import oracle as o  # oracle/oracle.py (the FP64 CPU oracle)
    o.lib()
    o.set_threads(min(os.cpu_count() or 1, 16))
    return o


@pytest.fixture(scope="session")
def gpu_ctx():
    import culamp_amd
    if culamp_amd.device_count() == 0:
        pytest.fail("no HIP device visible: the gpu tests need an MI355X")
    ctx = culamp_amd.Context(0)
    yield ctx
    ctx.close()
