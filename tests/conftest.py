import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libmmhip.so")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def M1():
    from micromagnum_amd.omf import read_omf
    return read_omf(os.path.join(GOLDEN, "M1.omf"))


def rel_err(a, b):
    """normwise relative error ||a-b||_2 / ||b||_2"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def max_rel_err(a, b):
    """max |a-b| / max |b|"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def random_unit_field(shape, seed, Ms=1.0):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((3, nz, ny, nx))
    m /= np.sqrt((m ** 2).sum(axis=0))
    return m * Ms


def vortex_field(shape, delta, core, polarization=1.0, chirality=1.0, core_radius=10e-9, Ms=8e5):
    """magnetizationFunction of src/magnum/micromagnetics/toolbox/vortex.py:24-44
    evaluated at cell centres (i+0.5)*delta (RectangularMesh.h:45-58)."""
    nx, ny, nz = shape
    dx, dy, dz = delta
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="xy")
    Mx = -(Y - core[1]) * chirality
    My = (X - core[0]) * chirality
    Mz = np.full_like(Mx, polarization * core_radius)
    s = Ms / np.sqrt(Mx ** 2 + My ** 2 + Mz ** 2)
    M = np.stack([Mx * s, My * s, Mz * s])
    return np.ascontiguousarray(np.broadcast_to(M[:, None, :, :], (3, nz, ny, nx)))


def gpu_available():
    """libmmhip.so built and its HIP runtime sees a device (torch is not
    initialised: the library brings its own HIP runtime)."""
    try:
        from micromagnum_amd.device import gpu_available as g
        return g()
    except Exception:
        return False
