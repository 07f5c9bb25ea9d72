"""The reference-style front end (micromagnum_amd/magnum.py) on the CPU:
world and material set-up, conditions and solver construction, which need
no GPU (stepping is tested in test_gpu_frontend.py)."""
import numpy as np
import pytest

from micromagnum_amd import magnum as mm


def _solver(modules, **kw):
    world = mm.World(mm.RectangularMesh((8, 6, 2), (2e-9, 3e-9, 4e-9)), mm.Material.Py(alpha=0.1))
    return mm.create_solver(world, modules, **kw)


def test_material_database_and_overrides():
    py = mm.Material.Py(alpha=0.02)
    assert (py.Ms, py.A, py.alpha) == (8e5, 13e-12, 0.02)  # material.py:111-126
    assert mm.Material.Co().k_uniaxial == 520e3
    with pytest.raises(AttributeError):
        py.k_cubic


def test_world_bodies_fill_parameters_per_cell():
    mesh = mm.RectangularMesh((10, 4, 1), (1e-9, 1e-9, 1e-9))
    world = mm.World(mesh, mm.Body("left", mm.Material.Py(), mm.Cuboid((0, 0, 0), (5e-9, 4e-9, 1e-9))),
                     mm.Body("right", mm.Material.Py(Ms=5e5, alpha=0.3), mm.Cuboid((5e-9, 0, 0), (10e-9, 4e-9, 1e-9))))
    s = mm.create_solver(world, [mm.ExchangeField])
    Ms, al = s.state.Ms, s.state.alpha
    assert np.all(Ms[0, :, :5] == 8e5) and np.all(Ms[0, :, 5:] == 5e5)
    assert np.all(al[0, :, :5] == 0.01) and np.all(al[0, :, 5:] == 0.3)
    assert np.all(s.state.A == 13e-12)
    with pytest.raises(ValueError):
        mm.World(mesh, mm.Material.Py(), mm.Body("x", mm.Material.Py()))


def test_create_solver_arguments():
    s = _solver([mm.StrayField, mm.ExchangeField])
    assert (s.method, s.eps_abs, s.eps_rel) == ("rkf45", 1e-3, 1e-4)  # create_solver.py:41-46
    assert _solver([], evolver="rk4").step_size == 3e-13
    assert _solver([], evolver="euler").step_size == 5e-15
    with pytest.raises(ValueError):
        _solver([], evolver="magic")
    with pytest.raises(ValueError):
        _solver([], epsabs=1e-3)


def test_parameters_and_magnetisation_assignment():
    s = _solver([mm.AnisotropyField, mm.ExternalField])
    s.state.k_uniaxial = 1e5
    s.state.axis1 = (0.0, 0.0, 1.0)
    s.state.H_ext_offs = (1e4, 0.0, 0.0)
    assert np.all(s.state.k_uniaxial == 1e5) and np.all(s.state.axis1[2] == 1.0)
    assert s.state.H_ext_offs == (1e4, 0.0, 0.0)
    # M is normalised to Ms with the reference's float norm when assigned
    s.state.M = (1.0, 2.0, 2.0)
    M = s._M
    e = np.float32(1.0) * np.float32(1.0) + np.float32(2.0) * np.float32(2.0) + np.float32(2.0) * np.float32(2.0)
    assert np.allclose(M[:, 0, 0, 0], np.array([1.0, 2.0, 2.0]) * (8e5 / float(np.sqrt(e))), rtol=0, atol=0)
    # a function of (field, position) at the cell centres
    s.state.M = lambda field, pos: (pos[0], field.mesh.size[1], 0.0)
    assert s._M[0, 1, 2, 3] > 0 and s._M[1, 0, 0, 0] > 0


class _S:
    def __init__(self, step, t=0.0, deg=None):
        self.step, self.t, self.deg_per_ns = step, t, deg


def test_conditions():
    c = mm.EveryNthStep(10)
    assert c.check(_S(20)) and not c.check(_S(21))
    assert c.steps_left(_S(20)) == 10 and c.steps_left(_S(27)) == 3
    t = mm.Time(1e-9)
    assert not t.check(_S(0, 5e-10)) and t.check(_S(0, 1e-9))
    both = mm.EveryNthStep(4) & mm.AfterNthStep(5)
    assert both.check(_S(8)) and not both.check(_S(4))
    assert both.steps_left(_S(6)) == 1
    assert (~mm.Never()).check(_S(3)) and (mm.Never() | mm.Always()).check(_S(3))
    e = mm.EveryNthSecond(1e-11)
    assert e.get_time_of_interest(_S(0, 2.5e-11)) == pytest.approx(3e-11)


def test_time_conditions_name_no_time_of_interest():
    # condition.py:145-167: TimeGreaterEq and TimeBetween inherit
    # get_time_of_interest() -> None, so solve(Time(t)) steps with
    # step_with_t_max(1e100) and its last step overshoots t; the front end
    # instead hands the device a stop time that does not clip
    t = mm.Time(1e-9)
    assert t.get_time_of_interest(_S(0, 5e-10)) is None
    assert t.stop_time(_S(0, 5e-10)) == 1e-9 and t.stop_time(_S(0, 1e-9)) is None
    assert t.steps_left(_S(0, 5e-10)) > 1000 and t.steps_left(_S(0, 2e-9)) == 1
    b = mm.TimeBetween(1e-9, 2e-9)
    assert b.get_time_of_interest(_S(0, 0.0)) is None and b.stop_time(_S(0, 0.0)) == 1e-9
    assert b.steps_left(_S(0, 1.5e-9)) == 1 and b.steps_left(_S(0, 3e-9)) > 1000
    assert (mm.Time(3e-9) | mm.Time(2e-9)).stop_time(_S(0, 0.0)) == 2e-9
    assert (mm.EveryNthStep(3) & mm.Time(2e-9)).stop_time(_S(0, 0.0)) == 2e-9
    assert (~mm.Time(2e-9)).stop_time(_S(0, 0.0)) == 2e-9
    assert mm.EveryNthStep(3).stop_time(_S(0, 0.0)) is None


def test_relaxed_condition_needs_a_previous_check_and_a_decrease():
    # condition.py:177-198
    r = mm.Relaxed(5.0, 3)
    st = _S(0, deg=1.0)
    assert not r.check(st)          # first check: nothing to compare with
    st.step, st.deg_per_ns = 3, 0.9
    assert r.check(st)              # decreasing and <= 5
    st.step, st.deg_per_ns = 4, 0.1
    assert not r.check(st)          # not a multiple of 3
    st.step, st.deg_per_ns = 6, 2.0
    assert not r.check(st)          # increased
