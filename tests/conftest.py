import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libspring_gpu.so")


@pytest.fixture(scope="session")
def rocksalt():
    from spring_amd import model
    return model.rocksalt2(with_fc4=True)


@pytest.fixture(scope="session")
def rocksalt10():
    from spring_amd import model
    return model.rocksalt10()


def problem(m, mesh_dims, T=300.0, time_rev=0, symmetry=True):
    from spring_amd import mesh, phonons
    qm = mesh.make_mesh(m.crystal, mesh_dims, time_rev=time_rev, symmetry=symmetry)
    ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, T)
    return qm, ph


def rel_err(a, b, floor_frac=1e-9):
    """max |a-b| / max(|b|, floor) with floor = floor_frac * max|b| (per array)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    floor = floor_frac * np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor + 1e-300)))
