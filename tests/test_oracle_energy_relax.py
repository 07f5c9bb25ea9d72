"""Pins of the oracle's energies and relaxation rule (CPU only).

Energies: closed forms of each term (Zeeman and anisotropy densities, the
discrete exchange energy of a periodic spiral, the demag energy of a single
cube, Nzz = 1/3).  Relaxation: the stop step and the stopped state are
recomputed from plain runs of the oracle with the rule evaluated in numpy
from its definition (condition.py:169-200, landau_lifshitz_gilbert.py:127-131).
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import random_unit_field
from micromagnum_amd import magneto as mg

MU0_PY = 4.0 * math.pi * 1e-7  # src/magnum/micromagnetics/constants.py:21


def _uniform(shape, v, Ms):
    nx, ny, nz = shape
    v = np.asarray(v, dtype=np.float64)
    v = v / np.linalg.norm(v)
    return np.ascontiguousarray(np.broadcast_to((Ms * v).reshape(3, 1, 1, 1), (3, nz, ny, nx)))


def test_zeeman_energy_closed_form():
    shape, delta, Ms, H0 = (4, 3, 2), (2e-9, 3e-9, 4e-9), 7e5, 3.3e4
    s = O.System(shape, delta, Ms=Ms, alpha=0.1, h_ext=(H0, 0.0, 0.0))
    V, n = np.prod(delta), np.prod(shape)
    E = s.energy(_uniform(shape, (1, 0, 0), Ms))
    assert E["E_ext"] == pytest.approx(-MU0_PY * V * n * Ms * H0, rel=1e-14)
    E = s.energy(_uniform(shape, (0, 1, 0), Ms))
    assert abs(E["E_ext"]) < 1e-30
    assert E["E_stray"] == 0.0 and E["E_exch"] == 0.0 and E["E_aniso"] == 0.0


@pytest.mark.parametrize("theta", [0.0, 0.4, math.pi / 2])
def test_uniaxial_energy_closed_form(theta):
    shape, delta, Ms, k = (3, 2, 2), (2e-9,) * 3, 8e5, 4.5e5
    s = O.System(shape, delta, Ms=Ms, alpha=0.1, k_uni=k, axis1=(0.0, 0.0, 1.0))
    E = s.energy(_uniform(shape, (math.sin(theta), 0.0, math.cos(theta)), Ms))
    want = k * math.sin(theta) ** 2 * np.prod(delta) * np.prod(shape)
    assert E["E_aniso"] == pytest.approx(want, rel=1e-13, abs=1e-40)
    assert E["E_tot"] == E["E_aniso"]


def test_cubic_energy_closed_form():
    shape, delta, Ms, k = (2, 2, 3), (2e-9,) * 3, 8e5, -5e3
    s = O.System(shape, delta, Ms=Ms, alpha=1.0, k_cub=k, axis1=(1.0, 0.0, 0.0), axis2=(0.0, 1.0, 0.0))
    V, n = np.prod(delta), np.prod(shape)
    E = s.energy(_uniform(shape, (1, 1, 1), Ms))   # a_i^2 = 1/3: k (3 * 1/9)
    assert E["E_aniso"] == pytest.approx(k / 3.0 * V * n, rel=1e-13)
    E = s.energy(_uniform(shape, (1, 1, 0), Ms))   # a1^2 a2^2 = 1/4
    assert E["E_aniso"] == pytest.approx(k / 4.0 * V * n, rel=1e-13)


def test_uniaxial_plus_cubic_energy_is_the_sum():
    shape, delta, Ms = (3, 3, 1), (2e-9,) * 3, 8e5
    M = random_unit_field(shape, 5, Ms)
    kw = dict(Ms=Ms, alpha=0.5, axis1=(0.0, 0.6, 0.8), axis2=(1.0, 0.0, 0.0))
    Eu = O.System(shape, delta, k_uni=3e5, **kw).energy(M)["E_aniso"]
    Ec = O.System(shape, delta, k_cub=2e4, **kw).energy(M)["E_aniso"]
    Eb = O.System(shape, delta, k_uni=3e5, k_cub=2e4, **kw).energy(M)["E_aniso"]
    assert Eb == pytest.approx(Eu + Ec, rel=1e-14)


@pytest.mark.parametrize("q", [1, 3])
def test_exchange_energy_of_periodic_spiral(q):
    # m_i = (cos(i t), sin(i t), 0), t = 2 pi q / nx, periodic in x: the
    # discrete Laplacian gives H = 2A/(mu0 Ms) (2 cos t - 2) m / dx^2, so
    # E = -mu0/2 V sum M.H = A (2 - 2 cos t) / dx^2 V N.
    nx, ny, nz = 12, 2, 1
    delta, Ms, A = (2e-9, 3e-9, 4e-9), 8e5, 1.3e-11
    t = 2 * math.pi * q / nx
    ang = np.arange(nx) * t
    M = np.zeros((3, nz, ny, nx))
    M[0] = Ms * np.cos(ang)
    M[1] = Ms * np.sin(ang)
    s = O.System((nx, ny, nz), delta, Ms=Ms, alpha=0.1, A=A, pbc=(True, False, False))
    E = s.energy(M)
    want = A * (2 - 2 * math.cos(t)) / delta[0] ** 2 * np.prod(delta) * nx * ny * nz
    assert E["E_exch"] == pytest.approx(want, rel=1e-12)


def test_demag_energy_of_a_cube():
    # one cubic cell, uniform M along z: H = -M/3, E = mu0/6 V Ms^2
    delta, Ms = (5e-9,) * 3, 8e5
    N = mg.GenerateDemagTensor(1, 1, 1, *delta)
    s = O.System((1, 1, 1), delta, Ms=Ms, alpha=0.1, N=N)
    E = s.energy(_uniform((1, 1, 1), (0, 0, 1), Ms))
    assert E["E_stray"] == pytest.approx(MU0_PY / 6.0 * np.prod(delta) * Ms ** 2, rel=1e-12)


def test_demag_energy_is_positive_and_shape_anisotropic():
    # a flat plate: in-plane magnetisation has less demag energy than out-of-plane
    shape, delta, Ms = (8, 8, 1), (3e-9, 3e-9, 1e-9), 8e5
    N = mg.GenerateDemagTensor(*shape, *delta)
    s = O.System(shape, delta, Ms=Ms, alpha=0.1, N=N)
    Ein = s.energy(_uniform(shape, (1, 0, 0), Ms))["E_stray"]
    Eout = s.energy(_uniform(shape, (0, 0, 1), Ms))["E_stray"]
    assert 0 < Ein < Eout < MU0_PY / 2 * Ms ** 2 * np.prod(delta) * np.prod(shape)


# ---------------------------------------------------------------- relaxation
def _relax_system():
    shape, delta = (8, 6, 1), (3e-9, 3e-9, 3e-9)
    N = mg.GenerateDemagTensor(*shape, *delta)
    s = O.System(shape, delta, Ms=8e5, alpha=1.0, A=1.3e-11, N=N, k_cub=-5e3,
                 axis1=(1.0, 0.0, 0.0), axis2=(0.0, 1.0, 0.0))
    return s, random_unit_field(shape, 11, 8e5)


def _scaled_dmdt_max(s, M, Ms=8e5):
    dM, _ = s.dmdt(M)
    return float(np.sqrt((dM ** 2).sum(axis=0)).max() / Ms)


def test_relax_dmdt_rule_stops_at_first_state_below_threshold():
    s, M0 = _relax_system()
    thr = 3e9
    M, info = s.relax(M0, method="rkf45", threshold=thr, criterion="dmdt", max_steps=400)
    assert info["relaxed"]
    k = info["steps"]
    assert k > 2
    # the stopped state is the state of a plain k-step run, and it satisfies the rule
    Mk, _ = s.run(M0, method="rkf45", steps=k)
    assert np.array_equal(M, Mk)
    assert _scaled_dmdt_max(s, M) < thr
    assert info["value"] == pytest.approx(_scaled_dmdt_max(s, M), rel=1e-14)
    # the state one step earlier does not
    Mk1, _ = s.run(M0, method="rkf45", steps=k - 1)
    assert _scaled_dmdt_max(s, Mk1) >= thr


def test_relax_dmdt_respects_max_steps():
    s, M0 = _relax_system()
    M, info = s.relax(M0, threshold=1.0, criterion="dmdt", max_steps=5)
    assert not info["relaxed"] and info["steps"] == 5
    M5, _ = s.run(M0, steps=5)
    assert np.array_equal(M, M5)


def test_relax_deg_per_ns_rule_matches_reference_condition():
    # Relaxed(max_degree_per_ns, check_every_nth_step) evaluated step by step
    s, M0 = _relax_system()
    thr, every = 2e4, 3
    M, info = s.relax(M0, method="rkf45", threshold=thr, criterion="deg_per_ns", check_every=every,
                      max_steps=300)
    assert info["relaxed"]
    last = None
    stop = None
    for step in range(0, 301):
        if step % every:
            continue
        if step == 0:
            Ms_, h = M0, 1e-13
        else:
            Ms_, st = s.run(M0, method="rkf45", steps=step)
            h = st["h"]
        dM, _ = s.dmdt(Ms_)
        dm = np.sqrt((dM ** 2).sum(axis=0)).max()
        mm = np.sqrt((Ms_ ** 2).sum(axis=0)).max()
        deg = 1e-9 * ((180.0 / math.pi) * math.atan2(dm * h, mm)) / h
        if last is not None and deg < last and deg <= thr:
            stop = step
            assert info["value"] == pytest.approx(deg, rel=1e-12)
            break
        last = deg
    assert stop is not None and info["steps"] == stop
