import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


def load_t(path):
    pos, tets = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "Vertex":
                pos.append([float(x) for x in parts[2:5]])
            elif parts[0] == "Tet":
                tets.append([int(x) for x in parts[2:6]])
    return np.asarray(pos), np.asarray(tets, dtype=np.int32)


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ctx():
    from gaia_amd import Context

    ctx = Context(0)
    yield ctx
    ctx.close()
