"""Pins of the oracle's step control, renormalisation, FSAL reuse and
relaxation rule against what arithmetic and the single-spin LLG fix.  CPU only.

Each check is chosen so that a plausible slip fails it: a float-vs-double
norm, a wrong step headroom or error exponent, a misplaced clamp or step
bound, a fresh instead of the carried-over k_0 of the FSAL methods, a
missing "decreasing" clause in the Relaxed condition.

Sources restated by the oracle:
  normalize3      src/magneto/matrix/device/cpu/CPUDevice.cpp:244-271
  controller      src/magnum/micromagnetics/micro_magnetics_stepsize_controller.py:35-73
  evolve / apply  src/magnum/evolver/runge_kutta.py:47-137
  Relaxed         src/magnum/solver/condition.py:169-200
  deg_per_ns      src/magnum/micromagnetics/landau_lifshitz_gilbert.py:127-131
"""
import math

import numpy as np
import pytest

import oracle as O

GAMMA = 2.211e5
MS = 8e5


# ------------------------------------------------------------ normalisation
def test_normalize_mode0_is_the_float_norm_of_float_components():
    # CPUDevice.cpp:255-259: e0..e2 are float, so std::sqrt(e0*e0+e1*e1+e2*e2)
    # is the float overload; numpy float32 arithmetic rounds every operation
    # to float in the same order, and np.sqrt of a float32 is correctly rounded
    rng = np.random.default_rng(0xA18)
    n = 4000
    M = rng.standard_normal((3, 1, 1, n)) * 10 ** rng.uniform(-3, 7, (1, 1, 1, n))
    M[:, 0, 0, :5] = 0.0
    M[0, 0, 0, 1] = 3.0  # exact norm 3
    M[:, 0, 0, 2] = (1e-30, 0.0, 0.0)  # underflows to 0 in float: left unchanged
    length = rng.uniform(1e5, 1.5e6, (1, 1, n))
    out = O.normalize(M, length, mode=0)
    e = M.astype(np.float32)
    s = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]
    norm = np.sqrt(s).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        want = np.where(norm == 0.0, M, M * (length / norm))
    assert np.array_equal(out, want)
    # the float norm differs from the double norm in most cells (so the two
    # modes are told apart), by a few float ulps (casts, sums, square root)
    dn = np.sqrt((M ** 2).sum(axis=0))
    mask = norm > 0
    rel = np.abs(norm[mask] - dn[mask]) / dn[mask]
    assert (rel > 0).mean() > 0.9 and rel.max() < 3.6e-7
    out1 = O.normalize(M, length, mode=1)
    assert not np.array_equal(out, out1)
    np.testing.assert_allclose(np.sqrt((out1[:, mask] ** 2).sum(axis=0)), length[mask], rtol=1e-15)


# ------------------------------------------------------------ controller
def _iterate_power_law(order, q, C, eps_abs, eps_rel, dydt, h0, n=80):
    """Drive the controller with an error that scales as C h^q; return the
    accepted ratios e_abs/eps_abs and e_rel/eps_rel at the fixed point."""
    h = h0
    for _ in range(n):
        e = C * h ** q
        acc, hn = O.controller(eps_abs, eps_rel, h, order, e, dydt)
        h = hn
    e = C * h ** q
    acc, hn = O.controller(eps_abs, eps_rel, h, order, e, dydt)
    assert acc and hn == pytest.approx(h, rel=1e-12)
    return e / eps_abs, e / (h * dydt) / eps_rel


@pytest.mark.parametrize("order,q", [(4, 5), (5, 5), (3, 3), (4, 4)])
def test_controller_fixed_point_ratios(order, q):
    # with e(h) = C h^q the step converges to h* with e(h*) = 0.85^(p+1) eps
    # when the absolute criterion binds, and e_rel(h*) = 0.85^p eps_rel when
    # the relative one binds (a headroom of 0.9 or exponents 1/p, 1/(p+1)
    # swapped give other ratios)
    # C puts the boundary step e(h_b) = eps at 1e-12 s, inside the step bounds
    ra, _ = _iterate_power_law(order, q, 1e-6 / 1e-12 ** q, 1e-6, 1e30, 1.0, 1e-13)
    assert ra == pytest.approx(0.85 ** (order + 1), rel=1e-9)
    _, rr = _iterate_power_law(order, q, 1e-4 * 1e9 / 1e-12 ** (q - 1), 1e30, 1e-4, 1e9, 1e-13)
    assert rr == pytest.approx(0.85 ** order, rel=1e-9)


def test_controller_accept_boundary_and_scale_clamps():
    eps, h, p = 1e-4, 1e-12, 4
    big = 1e20  # h * dydt large: the relative error is tiny
    assert O.controller(eps, eps, h, p, eps, big)[0]
    assert not O.controller(eps, eps, h, p, np.nextafter(eps, 1.0), big)[0]
    # relative error exactly at the bound: accepted; just above: rejected
    dydt = 1e9
    e_rel_at = eps * h * dydt
    assert O.controller(1.0, eps, h, p, e_rel_at, dydt)[0]
    assert not O.controller(1.0, eps, h, p, e_rel_at * (1 + 1e-12), dydt)[0]
    # a tiny error grows the step by the maximum scale 5, a huge one shrinks
    # it by the minimum scale 0.2
    acc, hn = O.controller(eps, eps, h, p, 1e-30, big)
    assert acc and hn == h * 5.0
    acc, hn = O.controller(eps, eps, h, p, 1e3, big)
    assert not acc and hn == h * 0.2


def test_controller_step_bounds_and_forced_accept():
    eps, p, big = 1e-4, 4, 1e20
    # never above 1e-10 s
    acc, hn = O.controller(eps, eps, 4e-11, p, 1e-30, big)
    assert acc and hn == 1e-10
    # never below 1e-15 s, and an attempt whose next step hits the floor is
    # accepted whatever its error
    acc, hn = O.controller(eps, eps, 3e-15, p, 1e3, big)
    assert acc and hn == 1e-15
    acc, hn = O.controller(eps, eps, 1e-15, p, 1e3, big)
    assert acc and hn == 1e-15
    # just above the floor after scaling: rejected
    acc, hn = O.controller(eps, eps, 5.1e-15, p, 1e3, big)
    assert not acc and hn == pytest.approx(5.1e-15 * 0.2, rel=1e-15) and hn > 1e-15


def test_controller_zero_denominator():
    # h * max_dydt == 0 raises ZeroDivisionError in the reference, which
    # returns (True, 5e-14) (micro_magnetics_stepsize_controller.py:45-50)
    assert O.controller(1e-4, 1e-4, 1e-12, 4, 1e-3, 0.0) == (True, 5e-14)
    assert O.controller(1e-4, 1e-4, 0.0, 4, 1e-3, 1e9) == (True, 5e-14)


# ------------------------------------------------------------ closed loop
def _precession_system(H0):
    # one cell, no damping: dM/dt = f1 M x H is a rotation about z, so every
    # step's error estimate is the same function of h alone
    return O.System((1, 1, 1), (2e-9,) * 3, Ms=MS, alpha=0.0, h_ext=(0.0, 0.0, H0), normalize_mode=1)


def _m0(theta=1.0, phi=0.3):
    return (MS * np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi),
                           np.cos(theta)])).reshape(3, 1, 1, 1)


def _first_attempt_accepted(s, method, h, eps_abs, eps_rel):
    _, st = s.run(_m0(), method=method, steps=1, eps_abs=eps_abs, eps_rel=eps_rel, h0=h)
    return st["rejections"] == 0


def _boundary_step(s, method, eps_abs, eps_rel, lo, hi):
    """Largest h whose first attempt the oracle accepts (bisection)."""
    assert _first_attempt_accepted(s, method, lo, eps_abs, eps_rel)
    assert not _first_attempt_accepted(s, method, hi, eps_abs, eps_rel)
    for _ in range(60):
        mid = math.sqrt(lo * hi)
        if _first_attempt_accepted(s, method, mid, eps_abs, eps_rel):
            lo = mid
        else:
            hi = mid
    return lo


# (method, order p of the tableau, exponent q of its error estimate)
_LOOP_CASES = [("rkf45", 4, 5), ("cc45", 4, 5), ("dp54", 5, 5), ("rk23", 3, 3)]


@pytest.mark.parametrize("method,p,q", _LOOP_CASES)
@pytest.mark.parametrize("binding", ["abs", "rel"])
def test_run_loop_step_size_fixed_point(method, p, q, binding):
    # RungeKutta.evolve + the controller on uniform precession: the accepted
    # step settles where e(h*) = 0.85^(p+1) eps_abs (absolute bound) or
    # e_rel(h*) = 0.85^p eps_rel (relative bound).  With e ~ h^q and the
    # boundary h_b (largest step the oracle accepts at once) e(h_b) = eps, so
    # h*/h_b = 0.85^((p+1)/q), resp. 0.85^(p/(q-1)).  A headroom of 0.9, the
    # tableau order off by one or the exponents swapped move the ratio by >=2 %.
    s = _precession_system(1e4)
    if binding == "abs":
        eps_abs, eps_rel, want = 1e-9, 1e30, 0.85 ** ((p + 1) / q)
    else:
        eps_abs, eps_rel, want = 1e30, 1e-8, 0.85 ** (p / (q - 1))
    hb = _boundary_step(s, method, eps_abs, eps_rel, 1e-15, 9e-11)
    _, st = s.run(_m0(), method=method, steps=120, eps_abs=eps_abs, eps_rel=eps_rel, h0=hb * 0.3)
    assert st["h_next"] == pytest.approx(st["h"], rel=2e-4)  # converged
    assert st["h_next"] / hb == pytest.approx(want, rel=4e-3), (st, hb)


# ------------------------------------------------------------ FSAL reuse
_DP_A = [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0]
_DP_B = [[], [1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
         [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
         [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
         [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]]
_DP_C5 = [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0]
_DP_C4 = [5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40]


def _spin_rhs(alpha, H):
    # LLG of one spin in a uniform field (llge_cpu.cpp:25-45 with the
    # factors of landau_lifshitz_gilbert.py:133-158)
    f1 = -GAMMA / (1 + alpha ** 2)
    f2 = -alpha * GAMMA / ((1 + alpha ** 2) * MS)

    def f(M):
        a = np.cross(M, H)
        return f1 * a + f2 * np.cross(M, a)
    return f


def _dp54_attempt(f, y, h, k0):
    """Dormand-Prince 5(4) attempt from y with the given k_0 (published
    coefficients): the 5th-order result, the error estimate, the last stage."""
    k = [k0]
    for s in range(1, 7):
        yt = y + h * sum(_DP_B[s][j] * k[j] for j in range(s))
        k.append(f(yt))
    y1 = y + h * sum(_DP_C5[j] * k[j] for j in range(7))
    err = h * sum((_DP_C4[j] - _DP_C5[j]) * k[j] for j in range(7))
    return y1, err, k[6]


def _dp54_evolve(f, y0, h0, eps, stale):
    """RungeKutta.evolve with dp54 on one spin: retries until the controller
    accepts.  stale=True: every retry takes k_0 from the last stage of the
    attempt before it (the reference); False: dM/dt at y each time."""
    k0 = f(y0)
    h, attempts = h0, 0
    while True:
        attempts += 1
        y1, err, k6 = _dp54_attempt(f, y0, h, k0)
        acc, hn = O.controller(eps, eps, h, 5, np.linalg.norm(err) / MS, np.linalg.norm(k0) / MS)
        if acc:
            return y1 * (MS / np.linalg.norm(y1)), attempts, h
        h = hn
        if stale:
            k0 = k6


def test_fsal_retry_starts_from_the_rejected_attempts_last_stage():
    # runge_kutta.py:97-101,131-134: state.dydt_in is overwritten by every
    # apply, so a retry after a rejection takes k_0 from the rejected
    # attempt's last stage (dM/dt at its end point), not dM/dt at y, and the
    # controller's relative error uses that k_0 too (apply returns k[0])
    alpha, H = 0.1, np.array([0.0, 0.0, 1e5])
    f = _spin_rhs(alpha, H)
    s = O.System((1, 1, 1), (2e-9,) * 3, Ms=MS, alpha=alpha, h_ext=tuple(H), normalize_mode=1)
    y0 = _m0().ravel()
    eps, h0 = 1e-7, 2e-11
    M, st = s.run(y0.reshape(3, 1, 1, 1), method="dp54", steps=1, eps_abs=eps, eps_rel=eps, h0=h0)
    y_st, n_st, h_st = _dp54_evolve(f, y0, h0, eps, stale=True)
    y_fr, n_fr, h_fr = _dp54_evolve(f, y0, h0, eps, stale=False)
    assert st["attempts"] == n_st and st["rejections"] == n_st - 1
    assert st["h"] == pytest.approx(h_st, rel=1e-12)
    assert np.abs(M.ravel() - y_st).max() / MS < 1e-13
    # a fresh k_0 per retry takes another path (fewer retries, other step)
    assert n_fr < n_st and abs(h_fr / h_st - 1) > 1e-3


def test_fsal_next_step_reuses_the_stage_before_renormalisation():
    # the accepted attempt's last stage is dM/dt at the unnormalised end
    # point; finish_step then renormalises M, and the next step still starts
    # from that carried-over k_0
    alpha, H = 0.3, np.array([2e4, 0.0, 1e5])
    f = _spin_rhs(alpha, H)
    s = O.System((1, 1, 1), (2e-9,) * 3, Ms=MS, alpha=alpha, h_ext=tuple(H), normalize_mode=1)
    y0 = _m0().ravel()
    eps, h0 = 1e-1, 1.5e-11
    M2, st2 = s.run(y0.reshape(3, 1, 1, 1), method="dp54", steps=2, eps_abs=eps, eps_rel=eps, h0=h0)
    _, st1 = s.run(y0.reshape(3, 1, 1, 1), method="dp54", steps=1, eps_abs=eps, eps_rel=eps, h0=h0)
    assert st2["rejections"] == 0 and st2["attempts"] == 2
    y1, _, k6 = _dp54_attempt(f, y0, st1["h"], f(y0))
    assert abs(np.linalg.norm(y1) / MS - 1) > 1e-8  # the step changed |M| measurably
    y1n = y1 * (MS / np.linalg.norm(y1))
    h2 = st2["t"] - st1["t"]
    carried, _, _ = _dp54_attempt(f, y1n, h2, k6)
    afresh, _, _ = _dp54_attempt(f, y1n, h2, f(y1n))
    carried *= MS / np.linalg.norm(carried)
    afresh *= MS / np.linalg.norm(afresh)
    d_ok = np.abs(M2.ravel() - carried).max() / MS
    d_alt = np.abs(M2.ravel() - afresh).max() / MS
    assert d_ok < 1e-14 and d_alt > 100 * max(d_ok, 1e-16)


# ------------------------------------------------------------ relaxation
def _damped_spin(alpha, H0):
    return O.System((1, 1, 1), (2e-9,) * 3, Ms=MS, alpha=alpha, h_ext=(0.0, 0.0, H0), normalize_mode=1)


def _analytic_theta(theta0, alpha, H0, t):
    g = GAMMA / (1 + alpha ** 2)
    return 2 * np.arctan(np.tan(theta0 / 2) * np.exp(-alpha * g * H0 * t))


def _analytic_scaled_dmdt(theta, alpha, H0):
    # |dM/dt| / Ms of a spin at angle theta to H: gamma H sin(theta) / sqrt(1+alpha^2)
    return GAMMA * H0 * np.sin(theta) / np.sqrt(1 + alpha ** 2)


def test_relax_dmdt_rule_against_the_analytic_damped_spin():
    # the run stops at the first accepted state where max|dM/dt|/Ms < thr,
    # which the closed-form theta(t) decides at the states' times
    alpha, H0, th0, thr = 0.5, 1e5, 1.0, 1e9
    s = _damped_spin(alpha, H0)
    M, info = s.relax(_m0(th0, 0.0), method="rkf45", threshold=thr, criterion="dmdt",
                      max_steps=5000, eps_abs=1e-9, eps_rel=1e-9)
    assert info["relaxed"] and info["steps"] > 5
    k = info["steps"]
    _, stk = s.run(_m0(th0, 0.0), method="rkf45", steps=k, eps_abs=1e-9, eps_rel=1e-9)
    _, stp = s.run(_m0(th0, 0.0), method="rkf45", steps=k - 1, eps_abs=1e-9, eps_rel=1e-9)
    vk = _analytic_scaled_dmdt(_analytic_theta(th0, alpha, H0, stk["t"]), alpha, H0)
    vp = _analytic_scaled_dmdt(_analytic_theta(th0, alpha, H0, stp["t"]), alpha, H0)
    assert vk < thr <= vp
    assert info["value"] == pytest.approx(vk, rel=1e-6)


def _deg_per_ns(v_scaled, h):
    # landau_lifshitz_gilbert.py:127-131 with max|M| = Ms
    return 1e-9 * (180.0 / math.pi) * math.atan2(v_scaled * MS * h, MS) / h


@pytest.mark.parametrize("thr,every", [(200.0, 4), (1e9, 3)])
def test_relax_deg_per_ns_rule_against_the_analytic_damped_spin(thr, every):
    # Relaxed(thr, every): checked at steps that are multiples of `every`;
    # relaxed iff deg_per_ns <= thr and below the previous check's value, so
    # the check at step 0 never stops the run (a huge thr stops it at the
    # second check, step `every`)
    alpha, H0, th0 = 0.5, 1e5, 1.0
    s = _damped_spin(alpha, H0)
    kw = dict(method="rkf45", eps_abs=1e-9, eps_rel=1e-9)
    M, info = s.relax(_m0(th0, 0.0), threshold=thr, criterion="deg_per_ns", check_every=every,
                      max_steps=5000, **kw)
    assert info["relaxed"]
    last, stop, val = None, None, None
    for step in range(0, info["steps"] + 1, every):
        if step == 0:
            t, h = 0.0, 1e-13
        else:
            _, st = s.run(_m0(th0, 0.0), steps=step, **kw)
            t, h = st["t"], st["h"]
        d = _deg_per_ns(_analytic_scaled_dmdt(_analytic_theta(th0, alpha, H0, t), alpha, H0), h)
        if last is not None and d < last and d <= thr:
            stop, val = step, d
            break
        last = d
    assert stop == info["steps"]
    assert info["value"] == pytest.approx(val, rel=1e-6)
    if thr >= 1e9:
        assert stop == every
