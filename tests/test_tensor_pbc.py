"""Periodic demag tensor fields (demag_tensor.cpp with periodic axes), CPU only.

Pins:
  * sum rules: summed over all entries, -N is the demag tensor of the
    infinite periodic body (rod along x: 0, 1/2, 1/2 for a square section;
    film in x-y: 0, 0, 1; a cube lattice: 1/3 each; trace 1 always);
  * the images summed cell by cell equal the open-boundary tensor of a long
    mesh summed over those images (the mirrored image ranges of
    tensor.cpp:51-64 and :167-180);
  * the far-image correction (tensor.cpp:66-160) agrees with summing the
    images explicitly much further out.
"""
import os

import numpy as np
import pytest

from micromagnum_amd import magneto as mg

D = (2e-9, 3e-9, 2.5e-9)


def tensor(shape, delta, pbc, repeat, correction=True):
    key = "MAGNUM_DEMAG_NO_INFINITY_CORRECTION"
    old = os.environ.pop(key, None)
    try:
        if not correction:
            os.environ[key] = "1"
        return mg.GenerateDemagTensor(*shape, *delta, *pbc, repeat=repeat)
    finally:
        os.environ.pop(key, None)
        if old is not None:
            os.environ[key] = old


def demag_sums(N):
    s = -N.reshape(-1, 6).sum(axis=0)
    return s[0], s[3], s[5], s[[1, 2, 4]]


def test_periodic_tensor_shape_is_mesh_shape():
    # periodic axes keep the mesh size, open ones are padded (round4(2n-1))
    N = tensor((6, 5, 3), D, (True, False, True), 2)
    assert N.shape == (3, 12, 6, 6)


def test_rod_sum_rule():
    xx, yy, zz, off = demag_sums(tensor((8, 1, 1), (2e-9, 2e-9, 2e-9), (True, False, False), 15))
    assert abs(xx) < 1e-9
    assert abs(yy - 0.5) < 1e-8 and abs(zz - 0.5) < 1e-8
    assert np.abs(off).max() < 1e-15


def test_rectangular_rod_trace_rule():
    # entries beyond 30 cells come from the truncated far-field series of the
    # open-boundary algorithm, whose trace is off by ~5e-9 per long axis (an
    # open 121x1x1 rod shows the same); the tail adds point dipoles
    xx, yy, zz, _ = demag_sums(tensor((8, 4, 1), D, (True, False, False), 15))
    assert abs(xx) < 2e-8
    assert abs(xx + yy + zz - 1.0) < 5e-8
    assert yy < zz  # thin direction z carries the larger factor


def test_film_sum_rule():
    xx, yy, zz, off = demag_sums(tensor((8, 8, 1), (2e-9, 2e-9, 2e-9), (True, True, False), 3))
    assert abs(zz - 1.0) < 1e-6 and abs(xx) < 1e-6 and abs(yy) < 1e-6
    # trace: exact for the tails, ~3e-8 from the far-field series of the images
    assert abs(xx + yy + zz - 1.0) < 1e-7


def test_cube_lattice_sum_rule():
    xx, yy, zz, off = demag_sums(tensor((4, 4, 4), (2e-9, 2e-9, 2e-9), (True, True, True), 2))
    for v in (xx, yy, zz):
        assert abs(v - 1.0 / 3.0) < 1e-12
    assert np.abs(off).max() < 1e-14


def test_correction_improves_the_sum_rule():
    shape, pbc = (8, 1, 1), (True, False, False)
    bare = demag_sums(tensor(shape, (2e-9,) * 3, pbc, 15, correction=False))
    corr = demag_sums(tensor(shape, (2e-9,) * 3, pbc, 15))
    assert abs(corr[1] - 0.5) < 1e-3 * abs(bare[1] - 0.5)


def _open_long(shape, delta, axis, reps):
    n = shape[axis]
    big = list(shape)
    big[axis] = reps * n + 1
    return tensor(tuple(big), delta, (False, False, False), 1)


def _open_entry(No, off):
    ez, ey, ex, _ = No.shape
    return No[off[2] % ez, off[1] % ey, off[0] % ex]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_explicit_images_equal_open_tensor_sums_1d(axis):
    shape, R = [5, 4, 3], 3
    pbc = [False, False, False]
    pbc[axis] = True
    Np = tensor(tuple(shape), D, tuple(pbc), R, correction=False)
    No = _open_long(shape, D, axis, R)
    n = shape[axis]
    worst = 0.0
    for K in range(-(shape[2] - 1), shape[2]):
        for J in range(-(shape[1] - 1), shape[1]):
            for I in range(-(shape[0] - 1), shape[0]):
                off = [I, J, K]
                if off[axis] < 0:
                    continue
                # images m in [-R, R-1] (the zero offset: [-(R-1), R-1])
                lo = -(R - 1) if off[axis] == 0 else -R
                acc = np.zeros(6)
                for m in range(lo, R):
                    o2 = list(off)
                    o2[axis] = off[axis] + m * n
                    acc += _open_entry(No, o2)
                idx = [I % (Np.shape[2]), J % Np.shape[1], K % Np.shape[0]]
                got = Np[idx[2], idx[1], idx[0]]
                worst = max(worst, np.abs(got - acc).max() / np.abs(No).max())
    assert worst < 1e-13, worst


def test_explicit_images_equal_open_tensor_sums_2d():
    shape, R = [4, 3, 2], 2
    Np = tensor(tuple(shape), D, (True, True, False), R, correction=False)
    big = (R * 4 + 1, R * 3 + 1, 2)
    No = tensor(big, D, (False, False, False), 1)
    worst = 0.0
    for K in range(-1, 2):
        for J in range(0, 3):
            for I in range(0, 4):
                acc = np.zeros(6)
                for m in range(-(R - 1) if I == 0 else -R, R):
                    for q in range(-(R - 1) if J == 0 else -R, R):
                        acc += _open_entry(No, [I + 4 * m, J + 3 * q, K])
                got = Np[K % Np.shape[0], J, I]
                worst = max(worst, np.abs(got - acc).max() / np.abs(No).max())
    assert worst < 1e-13, worst


def test_1d_correction_matches_far_explicit_images():
    # 15 images + tail estimate against 60 images + tail estimate; the tail
    # treats the far cells as point dipoles (relative size (l/x)^2 ~ 1e-4 of
    # the tail at 15 periods here)
    shape = (4, 3, 2)
    a = tensor(shape, D, (False, True, False), 15)
    b = tensor(shape, D, (False, True, False), 60)
    bare = tensor(shape, D, (False, True, False), 15, correction=False)
    err = np.abs(a - b).max()
    assert err < 2e-3 * np.abs(bare - b).max()
    assert err / np.abs(b).max() < 1e-8


def test_2d_correction_matches_far_explicit_images():
    # 3 images per axis + quadrant/strip tails against 12 images + tails; the
    # tail expansion is truncated at total degree 4, so agreement is limited
    shape = (4, 4, 2)
    a = tensor(shape, D, (True, False, True), 3)
    b = tensor(shape, D, (True, False, True), 12)
    bare = tensor(shape, D, (True, False, True), 3, correction=False)
    err = np.abs(a - b).max()
    assert err < 1e-2 * np.abs(bare - b).max()
    assert err / np.abs(b).max() < 1e-4
