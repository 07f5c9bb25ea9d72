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


def _film(mm, normalize_mode=0, **mat):
    world = mm.World(mm.RectangularMesh(SHAPE, DELTA), mm.Material.Py(**mat))
    solver = mm.create_solver(world, [mm.StrayField, mm.ExchangeField, mm.AnisotropyField, mm.ExternalField],
                              eps_abs=1e-4, eps_rel=1e-4, normalize_mode=normalize_mode)
    solver.state.k_uniaxial = 1e5
    solver.state.axis1 = (0.0, 0.0, 1.0)
    solver.state.H_ext_offs = (2e4, 0.0, 0.0)
    solver.state.M = vortex_field(SHAPE, DELTA, (48e-9, 32e-9), Ms=8e5)
    return solver


def _oracle(solver, alpha=0.1, normalize_mode=0):
    solver._stepper_now()  # the tensor the solver's stepper uses
    return O.System(SHAPE, DELTA, Ms=8e5, alpha=alpha, A=13e-12, N=solver._N, k_uni=1e5, axis1=(0.0, 0.0, 1.0),
                    h_ext=(2e4, 0.0, 0.0), normalize_mode=normalize_mode)


# The reference's renormalisation (normalize_mode 0) rounds |M| to float
# (CPUDevice.cpp:255-259).  Two FP64 paths whose sums round differently in
# the last double bit then, rarely, round a cell's norm to neighbouring
# floats: a 6e-8 relative jump of that cell, which the next steps spread to
# its neighbours, whose norms then land on other floats too.  Over tens of
# steps the states differ by up to ~1e-7 per cell while step counts, times
# and attempt sequences agree.  Long trajectories are therefore compared at
# 1e-11 with the FP64 norm (normalize_mode 1) and at 1e-6 per cell with the
# float norm.  The adaptive step control conditions trajectories further:
# the error estimate is a difference of nearly equal stage combinations, so
# a 1e-13 change of M moves the next step size by ~1e-10 relative; the
# oracle itself, started from M0 perturbed by 1e-16, ends 3e-11 away after
# the 55 steps of test_solve_time_with_step_handler.  Where a test passes
# `sens` (the oracle's own distance between runs from M0 and from M0 with a
# 1e-16 relative perturbation), the FP64 tolerance is max(1e-11, 10 sens).
M_TOL = {1: ("rel", 1e-11), 0: ("cell", 1e-6)}


def _close(M, Mo, mode, sens=0.0):
    kind, tol = M_TOL[mode]
    if kind == "rel":
        tol = max(tol, 10.0 * sens)
    err = rel_err(M, Mo) if kind == "rel" else np.abs(M - Mo).max() / 8e5
    print("normalize_mode %d: %s error %.3e (tolerance %.1e, oracle sensitivity %.1e)" % (mode, kind, err, tol,
                                                                                         sens))
    return err < tol


def _perturbed(M0, seed=1):
    return M0 * (1.0 + 1e-16 * np.random.default_rng(seed).standard_normal(M0.shape))


@pytest.mark.parametrize("mode", [1, 0])
def test_solve_time_with_step_handler(mm, mode):
    # solve(Time(t)) steps until state.t >= t without clipping the last step
    # (TimeGreaterEq names no time of interest, condition.py:145-157)
    solver = _film(mm, normalize_mode=mode, alpha=0.1)
    M0 = np.asarray(solver.state.M).copy()
    rec = Recorder(mm)
    solver.addStepHandler(rec, mm.EveryNthStep(5))
    t_end = 3e-11
    solver.solve(mm.Time(t_end))
    s = _oracle(solver, normalize_mode=mode)
    Mo, so = s.run_to_time(M0, t_end, eps_abs=1e-4, eps_rel=1e-4)
    Mp, sp = s.run_to_time(_perturbed(M0), t_end, eps_abs=1e-4, eps_rel=1e-4)
    assert solver.state.step == so["steps"] and solver.state.t == pytest.approx(so["t"], rel=1e-14)
    assert solver.state.t >= t_end
    assert _close(np.asarray(solver.state.M), Mo, mode, sens=rel_err(Mp, Mo) if sp["steps"] == so["steps"] else 0.0)
    # the handler saw exactly the states at steps 0, 5, 10, ...
    steps = [r[0] for r in rec.rows]
    assert steps == list(range(0, so["steps"] + 1, 5))
    for k, t, avg in rec.rows[1:4]:
        Mk, sk = s.run(M0, steps=k, eps_abs=1e-4, eps_rel=1e-4)
        assert t == pytest.approx(sk["t"], rel=1e-14)
        assert np.abs(avg - Mk.reshape(3, -1).mean(axis=1)).max() <= M_TOL[mode][1] * 8e5


def test_step_handler_time_of_interest_clips_steps(mm):
    # a handler condition with a time of interest (EveryNthSecond,
    # condition.py:121-133) clips the steps to its times: the stepping is
    # run_until(t_k) for t_k = dt, 2 dt, ... and the handler sees those states
    solver = _film(mm, alpha=0.1)
    M0 = np.asarray(solver.state.M).copy()
    rec = Recorder(mm)
    dt = 1e-11
    solver.addStepHandler(rec, mm.EveryNthSecond(dt))
    solver.solve(mm.Time(2 * dt))
    s = _oracle(solver)
    M1, s1 = s.run_until(M0, dt, eps_abs=1e-4, eps_rel=1e-4)
    assert [r[0] for r in rec.rows[:2]] == [0, s1["steps"]]
    assert rec.rows[1][1] == pytest.approx(dt, rel=1e-15)
    assert np.abs(rec.rows[1][2] - M1.reshape(3, -1).mean(axis=1)).max() <= 1e-11 * 8e5


@pytest.mark.parametrize("loop,mode", [("device", 1), ("solver", 1), ("device", 0)])
def test_relax_matches_the_relaxed_condition(mm, loop, mode):
    solver = _film(mm, normalize_mode=mode, alpha=0.5)
    M0 = np.asarray(solver.state.M).copy()
    if loop == "solver":  # a step handler sends relax() through solve()
        solver.addStepHandler(Recorder(mm), mm.EveryNthStep(1000))
    thr, every = 3e3, 10
    solver.relax(thr, every)
    s = _oracle(solver, alpha=0.5, normalize_mode=mode)
    Mo, so = s.relax(M0, criterion="deg_per_ns", threshold=thr, check_every=every, max_steps=20000)
    assert so["relaxed"] and solver.state.step == so["steps"]
    assert _close(np.asarray(solver.state.M), Mo, mode)


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
