"""mm_scaled_abs_max_noalloc (no counterpart in the reference: the variant of
mm_scaled_abs_max for callers that cannot allocate device memory) against
mm_scaled_abs_max, through the C ABI: two calls in a row, the second with a
smaller input, for a uniform and a per-cell scale.  The maximum and the
square root are exact or correctly rounded, so the results are equal."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]


def _reference(L, n, Md, Sd, uniform):
    out = ctypes.c_double()
    assert L.mm_scaled_abs_max(n, Md.ptr, Sd.ptr, int(uniform), ctypes.byref(out), None) == 0
    return out.value


def _noalloc(L, n, Md, Sd, uniform, outd):
    assert L.mm_scaled_abs_max_noalloc(n, Md.ptr, Sd.ptr, int(uniform), outd.ptr, None) == 0, L.mm_last_error()
    assert L.mm_synchronize(None) == 0
    return float(outd.numpy()[0])


@pytest.mark.parametrize("uniform", [True, False])
def test_noalloc_equals_scaled_abs_max_on_two_calls(uniform):
    from micromagnum_amd import _lib
    from micromagnum_amd.device import to_device, empty
    L = _lib.device_lib()
    rng = np.random.default_rng(11)
    outd = empty((1,))
    for n in (700001, 1531):  # the second call smaller: no stale partial maxima
        M = rng.standard_normal((3, n)) * 8e5
        S = np.full(n, 8e5) if uniform else rng.uniform(4e5, 1.2e6, n)
        if not uniform:
            S[::7] = 0.0  # cells with scale 0 are skipped
        Md, Sd = to_device(M), to_device(S)
        ref = _reference(L, n, Md, Sd, uniform)
        got = _noalloc(L, n, Md, Sd, uniform, outd)
        assert got == ref, (n, got, ref)
        # and both against numpy
        q = np.sqrt((M ** 2).sum(axis=0))
        want = q.max() / S[0] if uniform else (q[S != 0] / S[S != 0]).max()
        assert got == pytest.approx(want, rel=1e-15)


def test_noalloc_rejects_bad_arguments():
    from micromagnum_amd import _lib
    L = _lib.device_lib()
    assert L.mm_scaled_abs_max_noalloc(0, None, None, 1, None, None) != 0
    assert b"bad argument" in L.mm_last_error()
