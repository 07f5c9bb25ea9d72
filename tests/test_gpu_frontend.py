"""Reference-style scripts on the front end (micromagnum_amd/magnum.py)
against the oracle: create_solver + solve(Time) with a step handler,
relax() on the device and through the solver loop, a parameter change
between two solves, and the energies of the state.  FP64 at <= 1e-11 on M."""
import numpy as np
import pytest

import oracle as O
from conftest import gpu_available, rel_err, vortex_field

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


@pytest.fixture(scope="module")
def mm():
    from micromagnum_amd import magnum
    return magnum


SHAPE, DELTA = (24, 16, 2), (4e-9, 4e-9, 3e-9)


def Recorder(mm):
    """a StepHandler that keeps (step, t, <M>) of every state it handles"""
    class Rec(mm.StepHandler):
        def __init__(self):
            self.rows = []

        def handle(self, state):
            self.rows.append((state.step, state.t, np.asarray(state.M).reshape(3, -1).mean(axis=1)))
    return Rec()


def _film(mm, **mat):
    world = mm.World(mm.RectangularMesh(SHAPE, DELTA), mm.Material.Py(**mat))
    solver = mm.create_solver(world, [mm.StrayField, mm.ExchangeField, mm.AnisotropyField, mm.ExternalField],
                              eps_abs=1e-4, eps_rel=1e-4)
    solver.state.k_uniaxial = 1e5
    solver.state.axis1 = (0.0, 0.0, 1.0)
    solver.state.H_ext_offs = (2e4, 0.0, 0.0)
    solver.state.M = vortex_field(SHAPE, DELTA, (48e-9, 32e-9), Ms=8e5)
    return solver


def _oracle(solver, alpha=0.1):
    solver._stepper_now()  # the tensor the solver's stepper uses
    return O.System(SHAPE, DELTA, Ms=8e5, alpha=alpha, A=13e-12, N=solver._N, k_uni=1e5, axis1=(0.0, 0.0, 1.0),
                    h_ext=(2e4, 0.0, 0.0))


def test_solve_time_with_step_handler(mm):
    solver = _film(mm, alpha=0.1)
    M0 = np.asarray(solver.state.M).copy()
    rec = Recorder(mm)
    solver.addStepHandler(rec, mm.EveryNthStep(5))
    t_end = 3e-11
    solver.solve(mm.Time(t_end))
    s = _oracle(solver)
    Mo, so = s.run_until(M0, t_end, eps_abs=1e-4, eps_rel=1e-4)
    assert solver.state.step == so["steps"] and solver.state.t == pytest.approx(t_end, rel=1e-15)
    assert rel_err(np.asarray(solver.state.M), Mo) < 1e-11
    # the handler saw exactly the states at steps 0, 5, 10, ...
    steps = [r[0] for r in rec.rows]
    assert steps == list(range(0, so["steps"] + 1, 5))
    for k, t, avg in rec.rows[1:4]:
        if k == so["steps"]:
            continue
        Mk, sk = s.run(M0, steps=k, eps_abs=1e-4, eps_rel=1e-4)
        assert t == pytest.approx(sk["t"], rel=1e-14)
        assert np.abs(avg - Mk.reshape(3, -1).mean(axis=1)).max() <= 1e-11 * 8e5


@pytest.mark.parametrize("loop", ["device", "solver"])
def test_relax_matches_the_relaxed_condition(mm, loop):
    solver = _film(mm, alpha=0.5)
    M0 = np.asarray(solver.state.M).copy()
    if loop == "solver":  # a step handler sends relax() through solve()
        solver.addStepHandler(Recorder(mm), mm.EveryNthStep(1000))
    thr, every = 3e3, 10
    solver.relax(thr, every)
    s = _oracle(solver, alpha=0.5)
    Mo, so = s.relax(M0, criterion="deg_per_ns", threshold=thr, check_every=every, max_steps=20000)
    assert so["relaxed"] and solver.state.step == so["steps"]
    assert rel_err(np.asarray(solver.state.M), Mo) < 1e-11


def test_parameter_change_between_solves(mm):
    solver = _film(mm, alpha=0.1)
    M0 = np.asarray(solver.state.M).copy()
    solver.solve(mm.AfterNthStep(9))
    assert solver.state.step == 10
    solver.state.alpha = 0.4
    solver.solve(mm.Condition(lambda st: st.step >= 20))
    s1 = _oracle(solver, alpha=0.1)
    M1, st1 = s1.run(M0, steps=10, eps_abs=1e-4, eps_rel=1e-4)
    s2 = _oracle(solver, alpha=0.4)
    M2, st2 = s2.run(M1, steps=10, eps_abs=1e-4, eps_rel=1e-4, h0=st1["h_next"])
    assert solver.state.step == 20
    assert solver.state.t == pytest.approx(st1["t"] + st2["t"], rel=1e-14)
    assert rel_err(np.asarray(solver.state.M), M2) < 1e-11


def test_state_energies_and_fields(mm):
    solver = _film(mm, alpha=0.1)
    M = np.asarray(solver.state.M).copy()
    s = _oracle(solver)
    E = s.energy(M)
    for key in ("E_stray", "E_exch", "E_aniso", "E_ext", "E_tot"):
        assert getattr(solver.state, key) == pytest.approx(E[key], rel=1e-12, abs=1e-30)
    dM, H = s.dmdt(M)
    assert rel_err(np.asarray(solver.state.H_tot), H) < 1e-12
    assert rel_err(np.asarray(solver.state.dMdt), dM) < 1e-12
