"""Vortex helpers of the reference's toolbox
(src/magnum/micromagnetics/toolbox/vortex.py): the vortex magnetisation
function and the core finder, used for the core trajectory of the
spin-torque film workload.

findCore (vortex.py:80-134): the cell of largest |M_z| in the bottom plane,
refined to sub-cell precision by the vertex of the parabola through that
cell and its two neighbours along x and along y; the position is the cell
centre convention of RectangularMesh ((i + 0.5) * delta).
"""
import math

import numpy as np


def magnetizationFunction(core_x, core_y, polarization=+1, chirality=+1, core_radius=10e-9, Ms=8e5):
    """fn(field, pos) -> M of a vortex centred at (core_x, core_y)
    (vortex.py:24-44): in-plane curl, out-of-plane component
    polarization * core_radius before normalisation to Ms."""
    def fn(field, pos):
        x, y, z = pos
        mx = -(y - core_y) * chirality
        my = (x - core_x) * chirality
        mz = polarization * core_radius
        s = Ms / math.sqrt(mx * mx + my * my + mz * mz)
        return mx * s, my * s, mz * s
    return fn


def _vertex(f0, f1, f2):
    """abscissa of the vertex of the parabola through (-1, f0), (0, f1),
    (1, f2), relative to the middle point (0 when the points are collinear)"""
    den = f0 - 2.0 * f1 + f2
    return 0.0 if den == 0.0 else 0.5 * (f0 - f2) / den


def find_core(M, delta, origin=(0.0, 0.0)):
    """Core position (x, y) in metres of the vortex in M (3, nz, ny, nx):
    largest |M_z| in the plane z = 0, refined by a parabola through the
    neighbouring cells along x and y (no refinement along an axis where the
    maximum sits on the boundary)."""
    mz = np.abs(np.asarray(M)[2, 0])
    ny, nx = mz.shape
    j, i = np.unravel_index(int(np.argmax(mz)), mz.shape)
    fx = _vertex(mz[j, i - 1], mz[j, i], mz[j, i + 1]) if 0 < i < nx - 1 else 0.0
    fy = _vertex(mz[j - 1, i], mz[j, i], mz[j + 1, i]) if 0 < j < ny - 1 else 0.0
    return ((i + 0.5 + fx) * delta[0] - origin[0], (j + 0.5 + fy) * delta[1] - origin[1])


def findCore(solver, origin_x=0.0, origin_y=0.0):
    """The reference's findCore(solver, origin_x, origin_y) on a front-end
    solver (micromagnum_amd.magnum)."""
    return find_core(np.asarray(solver.state.M), solver.mesh.delta, (origin_x, origin_y))


def gaussian_core_vortex(shape, delta, center, width, Ms, polarization=+1, chirality=+1):
    """Vortex with an in-plane curl and a Gaussian core profile
    m_z = polarization exp(-r^2 / (2 width^2)), |m| = 1, times Ms, at the
    cell centres; (3, nz, ny, nx), the same in every plane."""
    nx, ny, nz = shape
    x = (np.arange(nx) + 0.5) * delta[0] - center[0]
    y = (np.arange(ny) + 0.5) * delta[1] - center[1]
    X, Y = np.meshgrid(x, y, indexing="xy")
    r = np.sqrt(X * X + Y * Y)
    mz = polarization * np.exp(-(r * r) / (2.0 * width * width))
    mp = np.sqrt(np.maximum(0.0, 1.0 - mz * mz))
    with np.errstate(invalid="ignore", divide="ignore"):
        cx = np.where(r > 0, -Y / r, 0.0) * chirality
        cy = np.where(r > 0, X / r, 0.0) * chirality
    M = np.stack([cx * mp, cy * mp, mz]) * Ms
    return np.ascontiguousarray(np.broadcast_to(M[:, None, :, :], (3, nz, ny, nx)))
