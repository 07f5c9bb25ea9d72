"""Vortex toolbox (micromagnum_amd/vortex.py) against closed forms: the core
finder of vortex.py:80-134 recovers the vertex of a sampled parabola exactly
and the centre of a symmetric Gaussian core; the Gaussian-core vortex has
unit |m|, the curl's sense and the core's profile."""
import numpy as np
import pytest

from micromagnum_amd.vortex import find_core, gaussian_core_vortex, magnetizationFunction


def test_find_core_recovers_a_parabola_vertex_exactly():
    # |M_z| is an exact paraboloid around (x0, y0) near its maximum: the
    # three-point vertex formula returns x0, y0 (cell centres at (i+0.5) d)
    nx, ny, d = 40, 30, (2e-9, 3e-9)
    x0, y0 = 17.3 * d[0], 11.8 * d[1]
    xc = (np.arange(nx) + 0.5) * d[0]
    yc = (np.arange(ny) + 0.5) * d[1]
    X, Y = np.meshgrid(xc, yc, indexing="xy")
    mz = 1.0 - ((X - x0) / 1e-7) ** 2 - ((Y - y0) / 1e-7) ** 2
    M = np.zeros((3, 2, ny, nx))
    M[2, 0] = mz
    M[2, 1] = -5.0  # only the bottom plane counts
    cx, cy = find_core(M, d)
    assert cx == pytest.approx(x0, rel=1e-12) and cy == pytest.approx(y0, rel=1e-12)
    ox, oy = find_core(M, d, origin=(1e-9, 2e-9))
    assert ox == pytest.approx(x0 - 1e-9, rel=1e-12) and oy == pytest.approx(y0 - 2e-9, rel=1e-12)


def test_find_core_uses_abs_mz_and_stays_on_boundary_cells():
    M = np.zeros((3, 1, 5, 6))
    M[2, 0, 0, 5] = -1.0  # corner: no neighbours on two sides
    assert find_core(M, (1.0, 1.0)) == (5.5, 0.5)


def test_gaussian_core_vortex_profile():
    shape, d, Ms, w = (64, 48, 2), (2e-9, 2e-9, 3e-9), 8.6e5, 10e-9
    c = (64e-9, 48e-9)  # a cell corner: the core sits between four cells
    M = gaussian_core_vortex(shape, d, c, w, Ms)
    assert np.allclose(np.sqrt((M ** 2).sum(axis=0)), Ms, rtol=1e-15)
    x = (np.arange(64) + 0.5) * d[0] - c[0]
    y = (np.arange(48) + 0.5) * d[1] - c[1]
    X, Y = np.meshgrid(x, y, indexing="xy")
    assert np.allclose(M[2, 0] / Ms, np.exp(-(X * X + Y * Y) / (2 * w * w)), rtol=1e-14, atol=0)
    # counter-clockwise curl (chirality +1): m_xy is perpendicular to r
    assert np.allclose(M[0, 0] * X + M[1, 0] * Y, 0.0, atol=1e-9 * Ms)
    assert (M[1, 0][X > 0] > 0).all() and (M[0, 0][Y > 0] < 0).all()
    cx, cy = find_core(M, d)
    assert abs(cx - c[0]) < 1e-3 * d[0] and abs(cy - c[1]) < 1e-3 * d[1]
    assert np.array_equal(M[:, 0], M[:, 1])


def test_magnetization_function_matches_the_reference_form():
    # vortex.py:24-44: (-(y-yc) c, (x-xc) c, p r) scaled to 8e5
    fn = magnetizationFunction(5e-9, 7e-9, polarization=-1, chirality=1, core_radius=10e-9)
    v = np.array(fn(None, (8e-9, 3e-9, 0.0)))
    raw = np.array([4e-9, 3e-9, -10e-9])
    assert np.allclose(v, raw / np.linalg.norm(raw) * 8e5, rtol=1e-15)
