import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (runs on the MI355X box)")


@pytest.fixture(scope="session")
def ctx():
    from emfa_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def sphere(subdiv, radius):
    """Product icosphere + product RWG list (the mesh every parity test uses)."""
    from emfa_amd import hip
    xyz, tri = hip.icosphere(subdiv, radius)
    rwg, nb = hip.rwg_build(tri)
    assert nb == 0
    return xyz, tri, rwg


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
