"""Compensated (TwoSum) FP64 sums on the GPU: <M>, dot sums and the term
energies are sums over cells of terms that can cancel.  The device carries
each sum as a pair s + e (Knuth's TwoSum, as the reference's accurate_sum,
src/magneto/math/accurate_sum.h:26-50), so the result is the correctly
rounded sum to a few ulp where a plain double sum loses most digits; its
second-order error, eps^2 times the sum of |terms| per term added, stays
below one ulp of the result for the data here (cancelling pairs of 1e8
among O(1) values).  math.fsum gives the correctly rounded sum on the host."""
import math

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


def _cancelling(n, seed, big=1e8):
    """n values whose plain double sum is dominated by rounding: pairs of
    +-big scattered among O(1) values."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    idx = rng.choice(n, size=n // 4 * 2, replace=False)
    h = len(idx) // 2
    v[idx[:h]] = big * (1.0 + rng.random(h))
    v[idx[h:]] = -v[idx[:h]] + rng.standard_normal(h)
    return v


def _ulps(a, b):
    return abs(a - b) / max(abs(b) * 2.0 ** -52, 5e-324)


@pytest.mark.parametrize("n", [1000, 300001])
def test_average_and_dot_sum_are_compensated(n):
    from micromagnum_amd import magneto as mg
    from micromagnum_amd.device import to_device
    M = np.stack([_cancelling(n, 1), _cancelling(n, 2), _cancelling(n, 3)])
    avg = mg.average(to_device(M))
    for c in range(3):
        exact = math.fsum(M[c]) / n
        plain = float(np.sum(M[c])) / n
        assert _ulps(avg[c], exact) <= 4, (c, avg[c], exact, plain)
    # the plain (pairwise) sum of the same data is far off
    assert all(_ulps(float(np.sum(M[c])) / n, math.fsum(M[c]) / n) > 1e3 for c in range(3))
    B = np.stack([np.ones(n), np.ones(n), np.ones(n)])
    d = mg.dot_sum(to_device(M), to_device(B))
    exact = math.fsum(np.concatenate([M[0], M[1], M[2]]))
    assert _ulps(d, exact) <= 4, (d, exact)


def test_stepper_average_and_zeeman_energy_are_compensated():
    # Zeeman energy -mu0 V sum_i M_i . H with a uniform H along x: the sum
    # over cells of M_x; <M> from the stepper's own reduction
    from micromagnum_amd.solver import Stepper
    shape, delta = (64, 48, 3), (2e-9, 2e-9, 2e-9)
    n = 64 * 48 * 3
    M = np.stack([_cancelling(n, 5), np.zeros(n), np.zeros(n)]).reshape(3, 3, 48, 64)
    H = 1234.5
    st = Stepper(shape, delta, Ms=8e5, alpha=0.1, demag=False, h_ext=(H, 0.0, 0.0))
    st.set_m(M)
    avg = st.average_m()
    exact = math.fsum(M[0].ravel())
    assert _ulps(avg[0] * n, exact) <= 4, (avg[0] * n, exact)
    E = st.energy()
    mu0 = 4.0 * math.pi * 1e-7
    V = delta[0] * delta[1] * delta[2]
    e_exact = -mu0 * V * math.fsum(M[0].ravel() * H)
    assert E["E_ext"] == pytest.approx(e_exact, rel=1e-14)
    st.close()
