"""GPU demag tensor generator (mm_generate_demag_tensor_gpu) against the host
long-double generator (mm_generate_demag_tensor, pinned by exact integrals
and closed forms in test_tensor.py), per branch:

  * entries evaluated on the host (finite-difference branches, marginal
    branch choices): bit-identical;
  * series / asymptotic entries, FP64 on the GPU: <= TOL_FAR relative to the
    largest entry of the component within the same shell of distance, and
    <= TOL_FAR_ENTRY relative to the entry itself where it is not tiny;
  * near field in FP64 on the GPU (near_on_gpu): cancellation grows as r^6
    in the distance r (cells); bound stated in test_near_field_in_fp64.
"""
import math

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]

TOL_FAR = 1e-13          # series / asymptotic, FP64 vs long double
TOL_FAR_ENTRY = 1e-11    # same, relative to the entry (entries above 1e-6 of the component's max)


@pytest.fixture(scope="module")
def mg():
    from micromagnum_amd import magneto
    return magneto


def _split(mg, shape, delta, near_on_gpu=False):
    info = {}
    Ng = mg.GenerateDemagTensor(*shape, *delta, device="gpu", near_on_gpu=near_on_gpu, info=info)
    Nh = mg.GenerateDemagTensor(*shape, *delta)
    return Ng, Nh, info


def _octant(N, shape):
    nx, ny, nz = shape
    return N[:nz, :ny, :nx, :]


@pytest.mark.parametrize("shape,delta", [
    ((72, 40, 4), (1e-9, 1e-9, 1e-9)),        # cubic cells
    ((128, 32, 1), (4e-9, 4e-9, 3e-9)),       # the thin-film workload's cells
    ((96, 64, 2), (3e-9, 3e-9, 2e-9)),        # the spin-torque film's cells
    ((160, 40, 3), (1e-9, 1.5e-9, 2e-9)),     # elongated cells: all three axes expanded only
                                              # beyond 60 cells (radius 30 x the longest edge)
])
def test_gpu_tensor_matches_host_generator(mg, shape, delta):
    Ng, Nh, info = _split(mg, shape, delta)
    assert Ng.shape == Nh.shape
    n = int(np.prod(shape))
    assert info["gpu_entries"] + info["host_entries"] == n
    assert info["gpu_entries"] > 0 and info["host_entries"] > 0
    # the padding gap is zero in both
    og, oh = _octant(Ng, shape), _octant(Nh, shape)
    eq = np.all(og == oh, axis=-1)
    assert eq.sum() >= info["host_entries"]
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((i * delta[0]) ** 2 + (j * delta[1]) ** 2 + (k * delta[2]) ** 2) / min(delta)
    worst_shell, worst_entry = 0.0, 0.0
    for lo in range(0, int(r.max()) + 1, 8):
        sel = (r >= lo) & (r < lo + 8)
        if not sel.any():
            continue
        for c in range(6):
            h, g = oh[..., c][sel], og[..., c][sel]
            scale = np.abs(h).max()
            if scale == 0:
                assert np.all(g == 0)
                continue
            worst_shell = max(worst_shell, np.abs(g - h).max() / scale)
            big = np.abs(h) > 1e-6 * np.abs(oh[..., c]).max()
            if big.any():
                worst_entry = max(worst_entry, (np.abs(g - h)[big] / np.abs(h)[big]).max())
    print("gpu tensor %s %s: gpu %d host %d entries, %.2f s + %.2f s; worst shell %.2e entry %.2e" % (
        shape, delta, info["gpu_entries"], info["host_entries"], info["gpu_s"], info["host_s"],
        worst_shell, worst_entry))
    assert worst_shell <= TOL_FAR and worst_entry <= TOL_FAR_ENTRY
    # the mirrored octants follow from the computed one exactly as on the host
    assert np.array_equal(Ng == 0, Nh == 0) or np.abs(Ng - Nh).max() <= TOL_FAR * np.abs(Nh).max()


def test_near_field_in_fp64(mg):
    # every entry on the GPU in FP64: the finite differences lose digits to
    # cancellation, ~eps * r^6 relative at distance r cells
    shape, delta = (40, 32, 6), (2e-9, 2e-9, 2e-9)
    Ng, Nh, info = _split(mg, shape, delta, near_on_gpu=True)
    assert info["host_entries"] == 0
    og, oh = _octant(Ng, shape), _octant(Nh, shape)
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt(i ** 2 + j ** 2 + k ** 2)
    rows = []
    worst = 0.0
    for lo in (0, 2, 4, 8, 16, 24, 30, 40):
        sel = (r >= lo) & (r < (lo * 2 if lo else 2))
        if not sel.any():
            continue
        e = 0.0
        for c in range(6):
            h, g = oh[..., c][sel], og[..., c][sel]
            s = np.abs(h).max()
            if s > 0:
                e = max(e, np.abs(g - h).max() / s)
        bound = 2.2e-16 * (max(lo, 1) * 2 + 2) ** 6
        rows.append((lo, e, bound))
        worst = max(worst, e / bound)
    print("near field in FP64 (r range, rel err, bound):", rows)
    assert worst <= 1.0


def test_large_grid_tensor_sampled_against_host_entries(mg):
    # the 512 x 512 x 64 grid of 2 nm cells (FFT 1024 x 1024 x 128): the GPU
    # generator against the host's long-double entries at sampled points
    shape, delta = (512, 512, 64), (2e-9, 2e-9, 2e-9)
    info = {}
    N = mg.GenerateDemagTensor(*shape, *delta, device="gpu", info=info)
    print("512x512x64 tensor: gpu %d host %d entries, %.2f s + %.2f s" % (
        info["gpu_entries"], info["host_entries"], info["gpu_s"], info["host_s"]))
    assert N.shape == (128, 1024, 1024, 6)
    rng = np.random.default_rng(0x7E2)
    pts = [(0, 0, 0), (1, 0, 0), (5, 3, 2), (29, 0, 0), (21, 21, 0), (31, 0, 0), (40, 10, 5),
           (511, 0, 0), (0, 511, 63), (300, 200, 63)]
    pts += [tuple(int(v) for v in (rng.integers(512), rng.integers(512), rng.integers(64))) for _ in range(150)]
    worst = 0.0
    for (i, j, k) in pts:
        h = mg.demag_tensor_entry(i, j, k, *delta)
        g = N[k, j, i]
        if math.sqrt(i * i + j * j + k * k) < 30:
            assert np.array_equal(g, h), (i, j, k, g, h)
        else:
            s = np.abs(h).max()
            worst = max(worst, np.abs(g - h).max() / s)
    assert worst <= TOL_FAR_ENTRY
    # self term: trace -1 (H = N M, a cube: -1/3 each)
    assert N[0, 0, 0, 0] + N[0, 0, 0, 3] + N[0, 0, 0, 5] == pytest.approx(-1.0, abs=1e-12)
    assert N[0, 0, 0, 0] == pytest.approx(-1.0 / 3.0, abs=1e-12)
    # far on the x axis the point dipole: Nxx = 2 V / (4 pi r^3)
    r = 500 * delta[0]
    V = delta[0] ** 3
    assert N[0, 0, 500, 0] == pytest.approx(2 * V / (4 * math.pi * r ** 3), rel=1e-4)
    # mirror symmetry: xy odd in x and y, xz odd in x and z
    assert N[3, 7, 1024 - 9, 1] == -N[3, 7, 9, 1]
    assert N[128 - 3, 7, 9, 2] == -N[3, 7, 9, 2]
