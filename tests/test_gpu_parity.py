"""GPU parity: every hot-path entry point of libmmhip.so against the CPU
oracle on the same seeded inputs.  FP64 bar: <= 1e-12 normwise relative
(rounding-order differences only); FP32 bar stated per test."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, gpu_available, rel_err, max_rel_err, random_unit_field, vortex_field

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")]

TOL64 = 1e-12
F64, F32 = np.float64, np.float32


def dev(a, dtype=F64):
    from micromagnum_amd.device import to_device
    return to_device(np.ascontiguousarray(a, dtype=np.float64), dtype)


def host(t):
    return t.numpy().astype(np.float64)


def like(t):
    from micromagnum_amd.device import empty_like
    return empty_like(t)


@pytest.fixture(scope="module")
def mg():
    from micromagnum_amd import magneto
    return magneto


def _omf(name):
    from micromagnum_amd.omf import read_omf
    return read_omf(os.path.join(GOLDEN, name))


# ---------------------------------------------------------------- exchange
def test_exchange_H1_fixture(mg):
    M = _omf("M1.omf")
    Href = _omf("H1_exch.omf")
    nx, ny, nz = M.num_nodes
    Ms = np.full(M.data.shape[1:], 8e5)
    A = np.full(M.data.shape[1:], 13e-12)
    H = like(dev(M.data))
    E = mg.exchange(nx, ny, nz, *M.delta, False, False, False, dev(Ms), dev(A), dev(M.data), H)
    Ho, Eo = O.exchange(M.delta, Ms, A, M.data)
    assert np.abs(host(H) - Href.data).max() < 2.0
    assert rel_err(host(H), Ho) < TOL64
    assert abs(E - Eo) <= TOL64 * abs(Eo)


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
@pytest.mark.parametrize("shape", [(13, 7, 5), (16, 16, 1), (1, 9, 4)])
def test_exchange_random_vs_oracle(mg, shape, pbc):
    nx, ny, nz = shape
    rng = np.random.default_rng(11)
    Ms = rng.uniform(4e5, 1.4e6, (nz, ny, nx))
    Ms[rng.random((nz, ny, nx)) < 0.1] = 0.0  # vacuum cells
    A = rng.uniform(5e-12, 2e-11, (nz, ny, nx))
    M = random_unit_field(shape, 12) * Ms
    delta = (2e-9, 3e-9, 2.5e-9)
    H = like(dev(M))
    E = mg.exchange(nx, ny, nz, *delta, *pbc, dev(Ms), dev(A), dev(M), H)
    Ho, Eo = O.exchange(delta, Ms, A, M, pbc)
    assert rel_err(host(H), Ho) < TOL64
    assert abs(E - Eo) <= 1e-11 * abs(Eo)


def test_exchange_f32_vs_oracle(mg):
    shape = (32, 24, 4)
    nx, ny, nz = shape
    Ms = np.full((nz, ny, nx), 8e5)
    A = np.full((nz, ny, nx), 1.3e-11)
    M = random_unit_field(shape, 3) * 8e5
    H = like(dev(M, F32))
    mg.exchange_f32(nx, ny, nz, 2e-9, 2e-9, 2e-9, 0, 0, 0, dev(Ms, F32), dev(A, F32),
                    dev(M, F32), H)
    Ho, _ = O.exchange((2e-9,) * 3, Ms, A, M)
    assert rel_err(host(H), Ho) < 1e-5  # FP32 bulk


# ---------------------------------------------------------------- anisotropy / llge / stt
def test_uniaxial_H1_fixture_and_random(mg):
    M = _omf("M1.omf")
    Href = _omf("H1_unianiso.omf")
    n = M.data[0].size
    axis = np.full(M.data.shape, 1.0 / np.sqrt(3.0))
    k = np.full(M.data.shape[1:], 520e3)
    Ms = np.full(M.data.shape[1:], 8e5)
    H = like(dev(M.data))
    E = mg.uniaxial_anisotropy(dev(axis), dev(k), dev(Ms), dev(M.data), H)
    Ho, Eo = O.uniaxial_anisotropy(axis, k, Ms, M.data)
    assert np.abs(host(H) - Href.data).max() < 10.0
    assert rel_err(host(H), Ho) < TOL64 and abs(E - Eo) <= TOL64 * abs(Eo)
    assert n == 10000


def test_anisotropy_random_vs_oracle(mg):
    shape = (17, 9, 3)
    nx, ny, nz = shape
    rng = np.random.default_rng(5)
    Ms = rng.uniform(4e5, 1.4e6, (nz, ny, nx))
    Ms[0, 0, :3] = 0.0
    k = rng.uniform(-1e6, 1e6, (nz, ny, nx))
    k[1, 2, :2] = 0.0
    M = random_unit_field(shape, 6) * Ms
    a1 = random_unit_field(shape, 7)
    tmp = random_unit_field(shape, 8)
    a2 = np.cross(a1, tmp, axis=0)
    a2 /= np.sqrt((a2 ** 2).sum(axis=0))
    H = like(dev(M))
    E = mg.uniaxial_anisotropy(dev(a1), dev(k), dev(Ms), dev(M), H)
    Ho, Eo = O.uniaxial_anisotropy(a1, k, Ms, M)
    assert rel_err(host(H), Ho) < TOL64 and abs(E - Eo) <= 1e-11 * abs(Eo)
    E = mg.cubic_anisotropy(dev(a1), dev(a2), dev(k), dev(Ms), dev(M), H)
    Ho, Eo = O.cubic_anisotropy(a1, a2, k, Ms, M)
    assert rel_err(host(H), Ho) < TOL64 and abs(E - Eo) <= 1e-11 * abs(Eo)


def test_llge_known_answer_and_random(mg):
    sh = (10, 10, 10)
    M = np.broadcast_to(np.array([5.0, 10.0, 15.0])[:, None, None, None], (3,) + sh)
    H = np.broadcast_to(np.array([20.0, 25.0, 30.0])[:, None, None, None], (3,) + sh)
    dM = like(dev(M))
    mg.llge(dev(np.full(sh, 10.0)), dev(np.full(sh, 20.0)), dev(M), dev(H), dM)
    d = host(dM)
    assert np.all(d[0] == -60750.0) and np.all(d[1] == -13500.0) and np.all(d[2] == 29250.0)
    rng = np.random.default_rng(9)
    f1, f2 = rng.standard_normal(sh), rng.standard_normal(sh)
    M, H = rng.standard_normal((3,) + sh), rng.standard_normal((3,) + sh)
    mg.llge(dev(f1), dev(f2), dev(M), dev(H), dM)
    assert rel_err(host(dM), O.llge(f1, f2, M, H)) < TOL64


@pytest.mark.parametrize("nz", [1, 20])
def test_zhangli_vortex_known_answer(mg, nz):
    shape = (100, 100, nz)
    delta = (3e-9, 3e-9, 3e-9)
    sh = (nz, 100, 100)
    M = vortex_field(shape, delta, (150e-9, 150e-9))
    j = np.broadcast_to(np.array([1e9, -1e9, 0.0])[:, None, None, None], (3,) + sh)
    args = [np.full(sh, 1.0), np.full(sh, 0.02), np.full(sh, 8e5), np.full(sh, 0.01), j, M]
    dM = like(dev(M))
    mg.fdm_zhangli(100, 100, nz, *delta, True, *[dev(a) for a in args], dM)
    avg = host(dM).reshape(3, -1).mean(axis=1)
    ref = (340104691300.73352, 337831145947.55255, 7.4024200439453129e-06)
    assert all(abs(a - b) < 1e6 for a, b in zip(avg, ref))
    assert rel_err(host(dM), O.zhangli(delta, True, *args)) < TOL64


def test_zhangli_random_vs_oracle(mg):
    shape = (12, 9, 4)
    nx, ny, nz = shape
    rng = np.random.default_rng(21)
    sh = (nz, ny, nx)
    Ms = rng.uniform(4e5, 1.4e6, sh)
    args = [rng.uniform(0.3, 1.0, sh), rng.uniform(0, 0.1, sh), Ms, rng.uniform(0.01, 0.5, sh),
            rng.standard_normal((3,) + sh) * 1e12, random_unit_field(shape, 22) * Ms]
    for prec in (True, False):
        dM = like(dev(args[-1]))
        mg.fdm_zhangli(nx, ny, nz, 2e-9, 3e-9, 4e-9, prec, *[dev(a) for a in args], dM)
        assert rel_err(host(dM), O.zhangli((2e-9, 3e-9, 4e-9), prec, *args)) < TOL64


# ---------------------------------------------------------------- reductions / normalise / RK
def test_reductions_vs_oracle(mg):
    shape = (40, 30, 7)
    nx, ny, nz = shape
    rng = np.random.default_rng(31)
    M = rng.standard_normal((3, nz, ny, nx)) * 1e5
    Ms = rng.uniform(4e5, 1.4e6, (nz, ny, nx))
    Ms[0, 0, 0] = 0.0
    assert abs(mg.scaled_abs_max(dev(M), dev(Ms)) - O.scaled_abs_max(M, Ms, uniform=False)) <= 1e-15
    Mu = np.full((nz, ny, nx), 8e5)
    assert mg.scaled_abs_max(dev(M), dev(Mu)) == O.scaled_abs_max(M, Mu, uniform=True)
    Mk = np.zeros((3, 10, 10, 10))
    Mk[0] = -8e5
    assert mg.scaled_abs_max(dev(Mk), dev(np.full((10, 10, 10), 8e5))) == 1.0
    B = rng.standard_normal(M.shape)
    ref = float(np.sum(M * B))
    assert abs(mg.dot_sum(dev(M), dev(B)) - ref) <= 1e-12 * np.sum(np.abs(M * B))
    avg = mg.average(dev(M))
    assert np.allclose(avg, M.reshape(3, -1).mean(axis=1), rtol=1e-12, atol=1e-9)
    assert abs(mg.abs_max(dev(M)) - np.sqrt((M ** 2).sum(axis=0)).max()) <= 1e-12 * 1e6


@pytest.mark.parametrize("mode", [0, 1])
def test_normalize_vs_oracle(mg, mode):
    shape = (33, 17, 3)
    rng = np.random.default_rng(41 + mode)
    M = rng.standard_normal((3, 3, 17, 33)) * 7e5
    M[:, 0, 0, 0] = 0.0  # zero vector kept
    Ms = rng.uniform(4e5, 1.4e6, (3, 17, 33))
    t = dev(M)
    mg.normalize(t, dev(Ms), mode)
    ref = O.normalize(M, Ms, mode)
    assert np.abs(host(t) - ref).max() <= 1e-15 * 1.4e6
    assert np.all(host(t)[:, 0, 0, 0] == 0.0)
    assert shape


@pytest.mark.parametrize("method", ["rkf45", "cc45", "dp54", "rk23", "rk4"])
def test_rk_algebra_vs_oracle(mg, method):
    tab = mg.ButcherTableau(method)
    ns = tab.getNumSteps()
    n = 300
    rng = np.random.default_rng(51)
    ks = [rng.standard_normal((3, 1, 1, n)) for _ in range(ns)]
    y = rng.standard_normal((3, 1, 1, n))
    h = 1.7e-13
    kd = [dev(k) for k in ks]
    ot = O.tableau(method)
    import ctypes
    karr = (ctypes.POINTER(ctypes.c_double) * 7)()
    kc = [np.ascontiguousarray(k.ravel()) for k in ks]
    for q in range(ns):
        karr[q] = kc[q].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t = O._Tableau()
    O.lib().orc_tableau_get(O.METHODS[method], ctypes.byref(t))
    yc = np.ascontiguousarray(y.ravel())
    for step in range(1, ns):
        yt = like(kd[0])
        mg.rk_prepare_step(step, h, tab, kd[:step] + [None] * (ns - step), dev(y), yt)
        ref = np.empty(3 * n)
        O.lib().orc_rk_prepare_step(step, ctypes.c_double(h), ctypes.byref(t), karr, O._p(yc), O._p(ref), n)
        assert rel_err(host(yt).ravel(), ref) < 1e-15
    yd = dev(y)
    ye = like(yd)
    mg.rk_combine_result(h, tab, kd, yd, ye)
    yref = yc.copy()
    eref = np.empty(3 * n)
    O.lib().orc_rk_combine_result(ctypes.c_double(h), ctypes.byref(t), karr, O._p(yref), O._p(eref), n)
    assert rel_err(host(yd).ravel(), yref) < 1e-15
    assert rel_err(host(ye).ravel(), eref) < 1e-12
    assert ot["num_steps"] == ns


# ---------------------------------------------------------------- demag convolution
def _conv_case(mg, shape, delta, precision, seed=1):
    nx, ny, nz = shape
    N = mg.GenerateDemagTensor(nx, ny, nz, *delta)
    M = random_unit_field(shape, seed) * 8e5
    conv = mg.SymmetricMatrixVectorConvolution_FFT(N, nx, ny, nz, precision=precision)
    dt = F64 if precision == "fp64" else F32
    H = like(dev(M, dt))
    conv.execute(dev(M, dt), H)
    Ho = O.Conv(N, nx, ny, nz).execute(M)
    return host(H), Ho, conv


def test_conv_H1_stray_fixture(mg):
    M = _omf("M1.omf")
    Hs = _omf("H1_stray.omf")
    nx, ny, nz = M.num_nodes
    N = mg.GenerateDemagTensor(nx, ny, nz, *M.delta)
    conv = mg.SymmetricMatrixVectorConvolution_FFT(N, nx, ny, nz)
    H = like(dev(M.data))
    conv.execute(dev(M.data), H)
    Ho = O.Conv(N, nx, ny, nz).execute(M.data)
    assert np.abs(host(H) - Hs.data).max() < 1.0  # stray_field_test.py:46-47
    assert rel_err(host(H), Ho) < TOL64


@pytest.mark.parametrize("shape,delta", [((24, 20, 6), (2e-9, 2e-9, 2e-9)), ((13, 11, 3), (3e-9, 2e-9, 4e-9)),
                                         ((64, 16, 1), (4e-9, 4e-9, 3e-9)), ((5, 40, 2), (2e-9, 2e-9, 2e-9)),
                                         ((1, 12, 7), (2e-9, 3e-9, 2e-9)), ((1, 1, 1), (2e-9, 2e-9, 2e-9))])
def test_conv_fp64_vs_oracle(mg, shape, delta):
    H, Ho, _ = _conv_case(mg, shape, delta, "fp64")
    assert rel_err(H, Ho) < TOL64, rel_err(H, Ho)
    assert max_rel_err(H, Ho) < 1e-11


@pytest.mark.parametrize("shape", [(24, 20, 6), (64, 16, 1), (13, 11, 3)])
def test_conv_fp32_vs_oracle(mg, shape):
    H, Ho, _ = _conv_case(mg, shape, (2e-9, 2e-9, 2e-9), "fp32")
    assert rel_err(H, Ho) < 1e-5, rel_err(H, Ho)  # FP32 FFT, FP32 accumulation


# padded lengths 4..2048 along every axis: the kernels' compile-time FFT plans
POW2_SHAPES = [(8, 8, 8), (16, 32, 4), (64, 128, 2), (256, 4, 3), (512, 2, 1), (1024, 1, 1), (3, 2, 1)]


@pytest.mark.parametrize("shape", POW2_SHAPES)
def test_conv_compiled_plans_vs_oracle_and_runtime_plans(mg, shape, monkeypatch):
    # power-of-two lengths run kernels instantiated for that length; the
    # runtime plans (forced) run the same passes in generic kernels
    H, Ho, _ = _conv_case(mg, shape, (2e-9, 3e-9, 2.5e-9), "fp64", seed=5)
    assert rel_err(H, Ho) < TOL64, rel_err(H, Ho)
    assert max_rel_err(H, Ho) < 1e-11
    monkeypatch.setenv("MM_FFT_GENERIC", "1")
    Hg, _, _ = _conv_case(mg, shape, (2e-9, 3e-9, 2.5e-9), "fp64", seed=5)
    assert rel_err(H, Hg) < 1e-14, rel_err(H, Hg)
    H32, _, _ = _conv_case(mg, shape, (2e-9, 3e-9, 2.5e-9), "fp32", seed=5)
    monkeypatch.delenv("MM_FFT_GENERIC")
    H32c, _, _ = _conv_case(mg, shape, (2e-9, 3e-9, 2.5e-9), "fp32", seed=5)
    assert rel_err(H32c, Ho) < 1e-5 and rel_err(H32, H32c) < 1e-6


@pytest.mark.parametrize("shape,pbc", [((1, 1, 1), (0, 0, 0)), ((3, 2, 1), (0, 0, 0)), ((4, 4, 1), (0, 0, 0)),
                                       ((2, 3, 5), (0, 0, 0)), ((5, 4, 3), (0, 0, 0)), ((17, 9, 2), (0, 0, 0)),
                                       ((6, 5, 2), (1, 0, 0))])
def test_conv_direct_sum_vs_oracle_and_fft(mg, shape, pbc):
    # SymmetricMatrixVectorConvolution_Simple: the direct sum over the mesh
    # (the reference's choice below 32 cells), against the oracle's direct sum
    # and the FFT convolution of the same tensor
    nx, ny, nz = shape
    delta = (2e-9, 3e-9, 2.5e-9)
    N = mg.GenerateDemagTensor(nx, ny, nz, *delta, *[bool(b) for b in pbc], repeat=2) if any(pbc) \
        else mg.GenerateDemagTensor(nx, ny, nz, *delta)
    M = random_unit_field(shape, 11) * 8e5
    conv = mg.SymmetricMatrixVectorConvolution_Simple(N, nx, ny, nz)
    H = like(dev(M))
    conv.execute(dev(M), H)
    Hs = host(H)
    Ho = O.conv_simple(N, M)
    assert rel_err(Hs, Ho) < 1e-13, rel_err(Hs, Ho)
    fft = mg.SymmetricMatrixVectorConvolution_FFT(N, nx, ny, nz)
    Hf = like(dev(M))
    fft.execute(dev(M), Hf)
    assert rel_err(Hs, host(Hf)) < 1e-12


@pytest.mark.parametrize("shape", [(24, 20, 6), (64, 16, 1), (13, 11, 3), (40, 36, 16)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_conv_wide_index_forms_give_identical_results(mg, shape, precision, monkeypatch):
    # the index arithmetic large grids take (integer division, 64-bit element
    # addressing; MM_WIDE_INDEX=1 forces it) against the 32-bit forms: the
    # same arithmetic on the same elements, so bit-identical fields
    H, Ho, _ = _conv_case(mg, shape, (2e-9, 2.5e-9, 3e-9), precision, seed=9)
    monkeypatch.setenv("MM_WIDE_INDEX", "1")
    Hw, _, _ = _conv_case(mg, shape, (2e-9, 2.5e-9, 3e-9), precision, seed=9)
    assert np.array_equal(H, Hw)
    assert rel_err(H, Ho) < (TOL64 if precision == "fp64" else 1e-5)


def test_conv_non_smooth_padding_reembedded(mg):
    # round4(2*13-1) = 28 = 4*7 is not radix-friendly: the library re-embeds the
    # tensor into 30; the linear convolution is unchanged.
    H, Ho, conv = _conv_case(mg, (13, 4, 1), (2e-9, 2e-9, 2e-9), "fp64")
    assert conv.fft_size[0] == 30
    assert rel_err(H, Ho) < TOL64


# ---------------------------------------------------------------- fused stepper
def _system(shape, with_stt=False, with_cubic=False):
    nx, ny, nz = shape
    kw = dict(Ms=1.1e6, alpha=0.02, A=1.5e-11, k_uni=5e5, axis1=(0.0, 0.0, 1.0), h_ext=(15915.49, 0.0, 0.0))
    if with_cubic:
        kw.update(k_cub=-5e3, axis2=(0.0, 1.0, 0.0), axis1=(1.0, 0.0, 0.0))
    if with_stt:
        kw.update(P=0.7, xi=0.05, j=(1e12, 0.0, 0.0))
    return kw


@pytest.mark.parametrize("shape,stt,cub", [((24, 20, 4), False, False), ((32, 32, 1), True, False),
                                           ((16, 12, 5), True, True)])
def test_stepper_dmdt_fp64_vs_oracle(shape, stt, cub):
    from micromagnum_amd.solver import Stepper
    kw = _system(shape, stt, cub)
    delta = (2e-9, 2e-9, 2e-9)
    st = Stepper(shape, delta, precision="fp64", **kw)
    M = random_unit_field(shape, 5) * 1.1e6
    st.set_m(M)
    dM, H = st.dmdt()
    sysO = O.System(shape, delta, N=st.N, **kw)
    dMo, Ho = sysO.dmdt(M)
    assert rel_err(H, Ho) < TOL64 and rel_err(dM, dMo) < TOL64
    st.close()


def test_stepper_dmdt_fp32_field_error():
    from micromagnum_amd.solver import Stepper
    shape = (48, 40, 8)
    kw = _system(shape)
    st = Stepper(shape, (2e-9,) * 3, precision="fp32", **kw)
    M = random_unit_field(shape, 6) * 1.1e6
    st.set_m(M)
    _, H = st.dmdt()
    Ho = O.System(shape, (2e-9,) * 3, N=st.N, **kw).h_tot(M)
    assert rel_err(H, Ho) < 1e-5, rel_err(H, Ho)
    st.close()


@pytest.mark.parametrize("method", ["rkf45", "cc45"])
def test_stepper_adaptive_trajectory_fp64_vs_oracle(method):
    from micromagnum_amd.solver import Stepper
    shape = (20, 16, 2)
    kw = _system(shape)
    st = Stepper(shape, (2e-9,) * 3, precision="fp64", method=method, eps_abs=1e-4, eps_rel=1e-4, **kw)
    M = random_unit_field(shape, 8) * 1.1e6
    st.set_m(M)
    st.run(25)
    Mg = st.get_m()
    s = st.stats()
    Mo, so = O.System(shape, (2e-9,) * 3, N=st.N, **kw).run(M, method=method, steps=25)
    assert s["attempts"] == so["attempts"] and s["rejections"] == so["rejections"]
    assert abs(s["t"] - so["t"]) <= 1e-12 * so["t"]
    assert np.abs(Mg - Mo).max() / 1.1e6 < 1e-10
    st.close()


def test_stepper_rk4_thin_film_vs_oracle():
    # the thin-film RK4 workload (128x32x1, 4x4x3 nm, Ms 8e5, A 1.3e-11, alpha 0.5)
    # for 100 of its 1000 steps; <M> recorded every 10 steps
    from micromagnum_amd.solver import Stepper
    shape, delta = (128, 32, 1), (4e-9, 4e-9, 3e-9)
    kw = dict(Ms=8e5, alpha=0.5, A=1.3e-11)
    st = Stepper(shape, delta, precision="fp64", method="rk4", h0=1e-13, **kw)
    M = random_unit_field(shape, 0x5EED) * 8e5
    st.set_m(M)
    avg = []
    for _ in range(10):
        st.run(10)
        avg.append(st.average_m())
    sysO = O.System(shape, delta, N=st.N, **kw)
    Mo, so = sysO.run(M, method="rk4", steps=100, h0=1e-13, record_avg=True)
    assert np.abs(st.get_m() - Mo).max() / 8e5 < 1e-10
    assert np.abs(np.array(avg) - so["avg"][9::10]).max() / 8e5 < 1e-10
    st.close()


def test_stepper_fp32_tracks_fp64():
    # both precisions stepped to the same end time (their step sequences differ)
    from micromagnum_amd.solver import Stepper
    shape = (32, 32, 4)
    kw = _system(shape)
    M = random_unit_field(shape, 10) * 1.1e6
    out = {}
    for prec in ("fp64", "fp32"):
        st = Stepper(shape, (2e-9,) * 3, precision=prec, method="rkf45", **kw)
        st.set_m(M)
        st.run_until(2e-12)
        out[prec] = (st.get_m(), st.stats())
        st.close()
    assert np.abs(out["fp32"][0] - out["fp64"][0]).max() / 1.1e6 < 1e-4


@pytest.mark.parametrize("method", ["dp54", "rk23"])
def test_stepper_fsal_trajectory_fp64_vs_oracle(method):
    # FSAL methods keep the reference's reuse of the last stage of the most
    # recent attempt (runge_kutta.py:97-101,131-134); the oracle does the same.
    # A spin-torque vortex: perturbing its start state by 1e-16 moves t and M
    # by < 1e-10 after 20 steps in the oracle (random states with rejected
    # attempts amplify such perturbations by up to 1e9, so they are no test).
    from micromagnum_amd.solver import Stepper
    shape, delta = (24, 20, 2), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, P=0.7, xi=0.05, j=(1e12, 0.0, 0.0))
    st = Stepper(shape, delta, precision="fp64", method=method, eps_abs=1e-5, eps_rel=1e-5, **kw)
    M = vortex_field(shape, delta, (36e-9, 30e-9), Ms=8.6e5)
    st.set_m(M)
    st.run(20)
    Mg = st.get_m()
    s = st.stats()
    Mo, so = O.System(shape, delta, N=st.N, **kw).run(M, method=method, steps=20, eps_abs=1e-5, eps_rel=1e-5)
    assert s["attempts"] == so["attempts"] and s["rejections"] == so["rejections"]
    assert abs(s["t"] - so["t"]) <= 1e-9 * so["t"]
    assert np.abs(Mg - Mo).max() / 8.6e5 < 1e-9
    st.close()


@pytest.mark.parametrize("method", ["rkf45", "dp54", "rk4"])
def test_stepper_alternating_field_vs_oracle(method):
    # Zeeman field offs + amp sin(t freq + phase) (alternating_field.py:57-84)
    # evaluated at every stage's time t + a_s h; then dM/dt and the energies
    # at the reached time
    from micromagnum_amd.solver import Stepper
    shape, delta = (20, 16, 2), (3e-9, 3e-9, 2e-9)
    ac = dict(h_ext=(1e4, 0.0, -5e3), h_ext_amp=(2e4, 3e4, 5e4), h_ext_freq=(2e10, 3e10, 5e10),
              h_ext_phase=(0.1, -0.7, 1.2))
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, **ac)
    # (h0 1e-13: no rejections, so DP54's carried-over k_0 keeps the
    # trajectory well conditioned; dM/dt after the run also checks that
    # evaluations outside an attempt ignore the finished run's stop flag)
    st = Stepper(shape, delta, precision="fp64", method=method, eps_abs=1e-5, eps_rel=1e-5, h0=1e-13, **kw)
    M = vortex_field(shape, delta, (30e-9, 24e-9), Ms=8.6e5)
    st.set_m(M)
    st.run(15)
    s = st.stats()
    sysO = O.System(shape, delta, N=st.N, **kw)
    Mo, so = sysO.run(M, method=method, steps=15, eps_abs=1e-5, eps_rel=1e-5, h0=1e-13)
    assert s["attempts"] == so["attempts"] and s["rejections"] == so["rejections"]
    assert abs(s["t"] - so["t"]) <= 1e-9 * so["t"]
    Mg = st.get_m()
    assert np.abs(Mg - Mo).max() / 8.6e5 < 1e-9
    dM, H = st.dmdt()
    dMo, Ho = sysO.dmdt(Mg, s["t"])
    assert rel_err(dM, dMo) < TOL64 and rel_err(H, Ho) < TOL64
    E, Eo = st.energy(), sysO.energy(Mg, s["t"])
    assert E["E_ext"] == pytest.approx(Eo["E_ext"], rel=1e-12)
    st.close()


def _pulse_field(t):
    # a static x field, a Gaussian pulse along y and a ramp along z
    import math
    g = math.exp(-((t - 4e-12) / 2e-12) ** 2)
    return (2e4, 6e4 * g, -3e3 + 1e15 * t)


@pytest.mark.parametrize("method", ["rkf45", "dp54", "rk4"])
def test_stepper_field_function_vs_oracle(method):
    # Zeeman field from a user function of t (alternating_field.py:60-75),
    # evaluated on the host at every stage time t + a_s h; then dM/dt and
    # the energies at the reached time
    from micromagnum_amd.solver import Stepper
    shape, delta = (20, 16, 2), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11)
    st = Stepper(shape, delta, precision="fp64", method=method, eps_abs=1e-5, eps_rel=1e-5, h0=1e-13,
                 h_ext_fn=_pulse_field, **kw)
    M = vortex_field(shape, delta, (30e-9, 24e-9), Ms=8.6e5)
    st.set_m(M)
    st.run(15)
    s = st.stats()
    sysO = O.System(shape, delta, N=st.N, h_ext_fn=_pulse_field, **kw)
    Mo, so = sysO.run(M, method=method, steps=15, eps_abs=1e-5, eps_rel=1e-5, h0=1e-13)
    assert s["attempts"] == so["attempts"] and s["rejections"] == so["rejections"]
    assert abs(s["t"] - so["t"]) <= 1e-9 * so["t"]
    Mg = st.get_m()
    assert np.abs(Mg - Mo).max() / 8.6e5 < 1e-9
    dM, H = st.dmdt()
    dMo, Ho = sysO.dmdt(Mg, s["t"])
    assert rel_err(dM, dMo) < TOL64 and rel_err(H, Ho) < TOL64
    E, Eo = st.energy(), sysO.energy(Mg, s["t"])
    assert E["E_ext"] == pytest.approx(Eo["E_ext"], rel=1e-12)
    st.close()


def test_stepper_field_function_run_until_and_precisions():
    # run_until with the last step clipped, FP64 against the oracle and the
    # mixed FP32 path against FP64
    from micromagnum_amd import magneto as mg
    from micromagnum_amd.solver import Stepper
    shape, delta = (24, 20, 2), (3e-9, 3e-9, 2e-9)
    N = mg.GenerateDemagTensor(*shape, *delta)
    phys = dict(Ms=8.6e5, alpha=0.05, A=1.3e-11)
    M = vortex_field(shape, delta, (36e-9, 30e-9), Ms=8.6e5)
    t_end = 7.3e-12
    out = {}
    for prec in ("fp64", "fp32"):
        st = Stepper(shape, delta, N=N, precision=prec, method="rkf45", eps_abs=1e-5, eps_rel=1e-5, h0=1e-13,
                     h_ext_fn=_pulse_field, **phys)
        st.set_m(M)
        st.run_until(t_end)
        out[prec] = (st.get_m(), st.stats())
        st.close()
    Mo, so = O.System(shape, delta, N=N, h_ext_fn=_pulse_field, **phys).run_until(
        M, t_end, method="rkf45", eps_abs=1e-5, eps_rel=1e-5, h0=1e-13)
    (M64, s64), (M32, s32) = out["fp64"], out["fp32"]
    assert s64["t"] == pytest.approx(t_end, rel=1e-15) and s32["t"] == pytest.approx(t_end, rel=1e-15)
    assert s64["attempts"] == so["attempts"]
    assert np.abs(M64 - Mo).max() / 8.6e5 < 1e-9
    assert np.abs(M32 - M64).max() / 8.6e5 < 1e-4
    assert np.abs(M64 - M).max() / 8.6e5 > 1e-3


def test_stepper_field_function_equals_alternating_parameters():
    # the alternating field as a host function gives the device-evaluated
    # alternating field's trajectory (the stage times and sines may differ in
    # the last bit between host and device)
    import math
    from micromagnum_amd.solver import Stepper
    offs, amp, w, phi = (1e4, 0.0, -5e3), (2e4, 3e4, 5e4), (2e10, 3e10, 5e10), (0.1, -0.7, 1.2)
    fn = lambda t: tuple(offs[c] + amp[c] * math.sin(t * w[c] + phi[c]) for c in range(3))
    shape, delta = (20, 16, 2), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, precision="fp64", method="dp54", eps_abs=1e-5, eps_rel=1e-5,
              h0=1e-13)
    M = vortex_field(shape, delta, (30e-9, 24e-9), Ms=8.6e5)
    a = Stepper(shape, delta, h_ext=offs, h_ext_amp=amp, h_ext_freq=w, h_ext_phase=phi, **kw)
    b = Stepper(shape, delta, h_ext_fn=fn, **kw)
    for st in (a, b):
        st.set_m(M)
        st.run(12)
    sa, sb = a.stats(), b.stats()
    assert sa["attempts"] == sb["attempts"] and sa["t"] == pytest.approx(sb["t"], rel=1e-13)
    assert np.abs(a.get_m() - b.get_m()).max() / 8.6e5 < 1e-12
    a.close()
    b.close()


def test_stepper_field_function_errors_and_removal():
    from micromagnum_amd.solver import Stepper
    shape, delta = (8, 8, 1), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, precision="fp64", method="rkf45", demag=False)
    M = vortex_field(shape, delta, (12e-9, 12e-9), Ms=8.6e5)

    def bad(t):
        raise ZeroDivisionError("from the field function")

    st = Stepper(shape, delta, h_ext_fn=bad, **kw)
    st.set_m(M)
    with pytest.raises(ZeroDivisionError):
        st.run(1)
    with pytest.raises(ValueError):
        Stepper(shape, delta, h_ext_fn=_pulse_field, h_ext=(1.0, 0.0, 0.0), **kw)
    st.close()
    # removing the function leaves a system without Zeeman field
    st = Stepper(shape, delta, h_ext_fn=_pulse_field, **kw)
    st.set_field_fn(None)
    st.set_m(M)
    plain = Stepper(shape, delta, **kw)
    plain.set_m(M)
    st.run(5)
    plain.run(5)
    assert np.array_equal(st.get_m(), plain.get_m())
    assert st.energy()["E_ext"] == 0.0
    st.close()
    plain.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stepper_euler_vs_oracle(prec):
    # euler.py:27-36: fixed step, y += h dM/dt, normalise; run_until ends at
    # the first multiple of h at or past t_end (no clipping)
    from micromagnum_amd.solver import Stepper
    shape, delta = (20, 16, 2), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, k_uni=1e5, axis1=(0.0, 0.0, 1.0), h_ext=(2e4, 0.0, 0.0))
    h = 5e-15
    st = Stepper(shape, delta, precision=prec, method="euler", h0=h, **kw)
    M = vortex_field(shape, delta, (30e-9, 24e-9), Ms=8.6e5)
    st.set_m(M)
    st.run(40)
    s40 = st.stats()
    sysO = O.System(shape, delta, N=st.N, **kw)
    Mo, so = sysO.run(M, method="euler", steps=40, h0=h)
    assert s40["attempts"] == so["attempts"] == 40 and s40["rejections"] == 0
    assert s40["t"] == pytest.approx(so["t"], rel=1e-14)
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert np.abs(st.get_m() - Mo).max() / 8.6e5 < tol
    t_end = s40["t"] + 2.5 * h
    st.run_until(t_end)
    su = st.stats()
    assert su["attempts"] == 43 and su["t"] == pytest.approx(43 * h, rel=1e-12)
    Mu, sou = sysO.run_until(Mo, t_end - so["t"], method="euler", h0=h)
    assert sou["steps"] == 3
    assert np.abs(st.get_m() - Mu).max() / 8.6e5 < tol
    st.close()


@pytest.mark.parametrize("method", ["rkf45", "dp54", "rk4"])
def test_stepper_run_until_vs_oracle(method):
    # stepping to a fixed end time: the last accepted step is redone with
    # h = t_end - t (runge_kutta.py:74-78)
    from micromagnum_amd.solver import Stepper
    shape, delta = (20, 16, 2), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, P=0.7, xi=0.05, j=(1e12, 0.0, 0.0))
    st = Stepper(shape, delta, precision="fp64", method=method, eps_abs=1e-5, eps_rel=1e-5, h0=1e-13, **kw)
    M = vortex_field(shape, delta, (30e-9, 24e-9), Ms=8.6e5)
    st.set_m(M)
    t_end = 2.345e-12
    st.run_until(t_end)
    s = st.stats()
    Mo, so = O.System(shape, delta, N=st.N, **kw).run_until(M, t_end, method=method, eps_abs=1e-5, eps_rel=1e-5,
                                                            h0=1e-13)
    assert s["t"] == pytest.approx(t_end, rel=1e-15) and so["t"] == pytest.approx(t_end, rel=1e-15)
    assert s["attempts"] == so["attempts"] and s["rejections"] == so["rejections"]
    assert abs(s["h"] - so["h"]) <= 1e-9 * so["h"]
    assert np.abs(st.get_m() - Mo).max() / 8.6e5 < 1e-9
    # stepping on continues with the controller's step size
    st.run(3)
    assert st.stats()["t"] > t_end
    st.close()


def test_stepper_vortex_spin_torque_dp54_fp32_tracks_fp64():
    # the spin-torque vortex workload (3x3x2 nm cells, Ms 8.6e5, A 1.3e-11,
    # alpha 0.1, j = 1e12 A/m^2 along x, P 0.7, beta 0.05, DP54 tol 1e-5) on a
    # 64x64x2 film: FP32 bulk against the FP64 path
    from micromagnum_amd.solver import Stepper
    shape, delta = (64, 64, 2), (3e-9, 3e-9, 2e-9)
    nx, ny, nz = shape
    M = vortex_field(shape, delta, (nx * delta[0] / 2, ny * delta[1] / 2), Ms=8.6e5)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, P=0.7, xi=0.05, j=(1e12, 0.0, 0.0))
    out = {}
    for prec in ("fp64", "fp32"):
        st = Stepper(shape, delta, precision=prec, method="dp54", eps_abs=1e-5, eps_rel=1e-5, **kw)
        st.set_m(M)
        st.run_until(4e-12)
        out[prec] = st.get_m()
        st.close()
    # same end time; the clipped last dp54 step reuses the full step's k_0 in
    # both runs (reference FSAL reuse), which bounds the agreement
    assert np.abs(out["fp32"] - out["fp64"]).max() / 8.6e5 < 1e-3


# ---------------------------------------------------------------- energies and relaxation
@pytest.mark.parametrize("shape,stt,cub", [((24, 20, 4), False, False), ((16, 12, 5), True, True)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_stepper_energy_vs_oracle(shape, stt, cub, prec, tol):
    from micromagnum_amd.solver import Stepper
    kw = _system(shape, stt, cub)
    delta = (2e-9, 2e-9, 2e-9)
    st = Stepper(shape, delta, precision=prec, **kw)
    M = random_unit_field(shape, 21) * 1.1e6
    st.set_m(M)
    E = st.energy()
    Eo = O.System(shape, delta, N=st.N, **kw).energy(M)
    for key in ("E_stray", "E_exch", "E_aniso", "E_ext", "E_tot"):
        assert abs(E[key] - Eo[key]) <= tol * max(abs(Eo[key]), 1e-30), (key, E[key], Eo[key])
    st.close()


def _relax_case():
    # cubic film (Ms 8e5, A 1.3e-11, K1 -5e3 along x,y, alpha 1) from a vortex
    shape, delta = (24, 16, 2), (3e-9, 3e-9, 3e-9)
    kw = dict(Ms=8e5, alpha=1.0, A=1.3e-11, k_cub=-5e3, axis1=(1.0, 0.0, 0.0), axis2=(0.0, 1.0, 0.0))
    return shape, delta, kw, vortex_field(shape, delta, (36e-9, 24e-9), Ms=8e5)


@pytest.mark.parametrize("criterion,thr,every", [("dmdt", 2.8e10, 1), ("deg_per_ns", 2e4, 3)])
@pytest.mark.parametrize("method", ["rkf45", "dp54"])
def test_stepper_relax_vs_oracle(criterion, thr, every, method):
    # the rule fires within the first ~12 steps, before the first rejected
    # attempt, where both paths are well conditioned: same stop step, state
    # and checked value
    from micromagnum_amd.solver import Stepper
    shape, delta, kw, M = _relax_case()
    st = Stepper(shape, delta, precision="fp64", method=method, **kw)
    st.set_m(M)
    info = st.relax(threshold=thr, criterion=criterion, check_every=every, max_steps=400)
    Mg = st.get_m()
    Mo, io = O.System(shape, delta, N=st.N, **kw).relax(M, method=method, threshold=thr, criterion=criterion,
                                                         check_every=every, max_steps=400)
    assert io["relaxed"] and info["relaxed"]
    assert info["steps"] == io["steps"]
    assert abs(info["value"] - io["value"]) <= 1e-9 * io["value"]
    assert np.abs(Mg - Mo).max() / 8e5 < 1e-10
    # stepping on after a relaxation continues from the kept state
    a0 = st.stats()["attempts"]
    st.run(2)
    assert st.stats()["attempts"] >= a0 + 2 and np.isfinite(st.get_m()).all()
    st.close()


def test_stepper_long_relaxation_vs_oracle():
    # ~130 steps with rejected attempts: perturbing the start state by 1e-16
    # moves the oracle's final state by ~1e-5, so the paths are compared by
    # what the rule guarantees and by the energy of the relaxed states
    from micromagnum_amd.solver import Stepper
    shape, delta, kw, M = _relax_case()
    st = Stepper(shape, delta, precision="fp64", method="rkf45", **kw)
    st.set_m(M)
    info = st.relax(threshold=3e9, criterion="dmdt", max_steps=1000)
    E = st.energy()["E_tot"]
    sysO = O.System(shape, delta, N=st.N, **kw)
    Mo, io = sysO.relax(M, threshold=3e9, criterion="dmdt", max_steps=1000)
    assert info["relaxed"] and io["relaxed"]
    assert info["value"] < 3e9 and abs(info["steps"] - io["steps"]) <= 0.1 * io["steps"]
    assert abs(E - sysO.energy(Mo)["E_tot"]) <= 1e-4 * abs(sysO.energy(Mo)["E_tot"])
    st.close()


# ---------------------------------------------------------------- periodic boundaries
@pytest.mark.parametrize("shape,pbc", [((16, 12, 3), (True, False, False)), ((20, 15, 1), (True, True, False)),
                                       ((8, 9, 10), (True, True, True)), ((15, 8, 4), (False, False, True)),
                                       ((6, 5, 1), (False, True, False))])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_conv_periodic_vs_oracle(mg, shape, pbc, precision):
    # periodic axes: circular convolution of period n (odd and even n)
    nx, ny, nz = shape
    N = mg.GenerateDemagTensor(nx, ny, nz, 2e-9, 3e-9, 2.5e-9, *pbc, repeat=2)
    M = random_unit_field(shape, 41) * 8e5
    conv = mg.SymmetricMatrixVectorConvolution_FFT(N, nx, ny, nz, precision=precision)
    assert tuple(conv.fft_size[a] for a in range(3) if pbc[a]) == tuple(shape[a] for a in range(3) if pbc[a])
    dt = F64 if precision == "fp64" else F32
    H = like(dev(M, dt))
    conv.execute(dev(M, dt), H)
    Ho = O.Conv(N, nx, ny, nz).execute(M)
    if nx * ny * nz <= 800:
        assert rel_err(Ho, O.conv_simple(N, M)) < 1e-13
    tol = TOL64 if precision == "fp64" else 1e-5
    assert rel_err(host(H), Ho) < tol, rel_err(host(H), Ho)


def test_conv_periodic_rejects_non_smooth_axis(mg):
    N = mg.GenerateDemagTensor(7, 4, 1, 2e-9, 2e-9, 2e-9, True, False, False, repeat=1)
    with pytest.raises(RuntimeError):
        mg.SymmetricMatrixVectorConvolution_FFT(N, 7, 4, 1)


@pytest.mark.parametrize("pbc,method", [((True, False, False), "rkf45"), ((True, True, False), "dp54")])
def test_stepper_periodic_vs_oracle(pbc, method):
    # periodic exchange neighbours + periodic demag tensor in the fused stepper
    from micromagnum_amd.solver import Stepper
    shape, delta = (16, 10, 2), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8e5, alpha=0.3, A=1.3e-11, k_uni=2e5, axis1=(0.0, 0.0, 1.0), h_ext=(1e4, 0.0, 0.0))
    st = Stepper(shape, delta, precision="fp64", method=method, pbc=pbc, periodic_repeat=2, **kw)
    M = random_unit_field(shape, 43) * 8e5
    st.set_m(M)
    dM, H = st.dmdt()
    sysO = O.System(shape, delta, N=st.N, pbc=pbc, **kw)
    dMo, Ho = sysO.dmdt(M)
    assert rel_err(H, Ho) < TOL64 and rel_err(dM, dMo) < TOL64
    st.run(15)
    Mo, so = sysO.run(M, method=method, steps=15)
    s = st.stats()
    assert s["attempts"] == so["attempts"] and s["rejections"] == so["rejections"]
    assert np.abs(st.get_m() - Mo).max() / 8e5 < 1e-10
    st.close()


@pytest.mark.parametrize("per_graph", ["4", "16"])
@pytest.mark.parametrize("method", ["rkf45", "dp54", "rk4"])
def test_stepper_graph_replays_equal_eager_launches(method, per_graph, monkeypatch):
    # attempts replayed from captured HIP graphs (per_graph attempts to a
    # long graph, the rest one to a graph; attempts after the run has stopped
    # do nothing) against the same attempts launched kernel by kernel:
    # bitwise equal after every call, with eager work (dM/dt evaluation,
    # <M>) and synchronisation between the replays
    from micromagnum_amd.solver import Stepper
    shape, delta = (20, 16, 3), (3e-9, 3e-9, 2e-9)
    kw = _system(shape, with_stt=True)
    M = random_unit_field(shape, 51) * 1.1e6
    out = {}
    monkeypatch.setenv("MM_GRAPH_ATTEMPTS", per_graph)
    for graphs in ("1", "0"):
        monkeypatch.setenv("MM_STEPPER_GRAPHS", graphs)
        st = Stepper(shape, delta, precision="fp64", method=method, eps_abs=1e-4, eps_rel=1e-4, h0=5e-14, **kw)
        st.set_m(M)
        states = []
        for rep in range(4):
            st.run(3 + 7 * rep)
            st.sync()
            st.dmdt()
            states.append((st.get_m(), st.stats(), st.average_m()))
        st.set_m(M)  # back to a first attempt that computes k_0
        st.run(2)
        states.append((st.get_m(), st.stats(), st.average_m()))
        out[graphs] = states
        st.close()
    for a, b in zip(out["1"], out["0"]):
        assert np.array_equal(a[0], b[0])
        assert a[1] == b[1]
        assert np.array_equal(a[2], b[2])


# ---------------------------------------------------------------- ensembles
def _ensemble_specs(count, seed, precision="fp32", max_xy=96):
    # the ensemble workload's distributions (sizes U[32,max_xy]^2 x U[1,4],
    # Ms U[4e5,1.4e6], A U[5e-12,2e-11], K U[0,1e6] along a random axis,
    # cells U[2,5] nm, alpha 0.5, RKF45), small sizes for a test
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        shape = (int(rng.integers(32, max_xy + 1)), int(rng.integers(32, max_xy + 1)), int(rng.integers(1, 5)))
        delta = tuple(float(v) for v in rng.uniform(2e-9, 5e-9, 3))
        ax = rng.standard_normal(3)
        ax /= np.linalg.norm(ax)
        specs.append(dict(shape=shape, delta=delta, Ms=float(rng.uniform(4e5, 1.4e6)), alpha=0.5,
                          A=float(rng.uniform(5e-12, 2e-11)), k_uni=float(rng.uniform(0, 1e6)), axis1=tuple(ax),
                          method="rkf45", precision=precision))
    return specs


@pytest.mark.parametrize("batched", ["1", "0"])
def test_ensemble_equals_single_steppers_and_oracle(batched, monkeypatch):
    # batched: every pass of an attempt is one launch for all systems of a
    # precision (runtime FFT plans, so the single steppers compared with
    # take the runtime plans too); "0": each system on its own graphs
    from micromagnum_amd.solver import Ensemble, Stepper
    monkeypatch.setenv("MM_BATCH_LAUNCH", batched)
    specs = _ensemble_specs(6, 61)
    specs[0]["precision"] = "fp64"
    specs[3]["shape"] = (40, 36, 1)  # a 2-D member
    ens = Ensemble(specs, nstreams=3)
    Ms = []
    for i, sp in enumerate(specs):
        M = random_unit_field(sp["shape"], 100 + i) * sp["Ms"]
        Ms.append(M)
        ens.set_m(i, M)
    ens.run(12)
    monkeypatch.setenv("MM_FFT_GENERIC", batched)
    for i, sp in enumerate(specs):
        kw = {k: v for k, v in sp.items() if k not in ("shape", "delta", "Ms", "alpha")}
        st = Stepper(sp["shape"], sp["delta"], sp["Ms"], sp["alpha"], N=ens.specs[i].N, **kw)
        st.set_m(Ms[i])
        st.run(12)
        assert np.array_equal(ens.get_m(i), st.get_m()), i
        assert ens.stats(i) == st.stats(), i
        st.close()
    monkeypatch.delenv("MM_FFT_GENERIC")
    # the FP64 member against the oracle
    sp = specs[0]
    kw = {k: v for k, v in sp.items() if k in ("A", "k_uni", "axis1")}
    Mo, so = O.System(sp["shape"], sp["delta"], Ms=sp["Ms"], alpha=sp["alpha"], N=ens.specs[0].N, **kw).run(
        Ms[0], method="rkf45", steps=12)
    s = ens.stats(0)
    assert s["attempts"] == so["attempts"] and s["rejections"] == so["rejections"]
    assert np.abs(ens.get_m(0) - Mo).max() / sp["Ms"] < 1e-10
    # a second run continues from where the first stopped
    ens.run(3)
    assert all(ens.stats(i)["attempts"] - ens.stats(i)["rejections"] == 15 for i in range(len(specs)))
    ens.close()


@pytest.mark.parametrize("shape,method", [((32, 16, 2), "rkf45"), ((64, 8, 4), "rkf45"), ((1024, 4, 2), "dp54"),
                                          ((1024, 4, 2), "rkf45")])
def test_stepper_compiled_plans_match_runtime_plans_and_oracle(shape, method, monkeypatch):
    # padded lengths (64, 32, 4), (128, 16, 8) and (2048, 8, 4): the stage
    # kernel's x transforms and every pass of the field run compile-time plans
    from micromagnum_amd.solver import Stepper
    delta = (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, P=0.7, xi=0.05, j=(1e12, 0.0, 0.0))
    M = vortex_field(shape, delta, (shape[0] * 1.5e-9, shape[1] * 1.5e-9), Ms=8.6e5)
    out = {}
    for generic in ("0", "1"):
        monkeypatch.setenv("MM_FFT_GENERIC", generic)
        st = Stepper(shape, delta, precision="fp64", method=method, eps_abs=1e-5, eps_rel=1e-5, **kw)
        st.set_m(M)
        st.run(10)
        out[generic] = (st.get_m(), st.stats(), st.energy(), st.N)
        st.close()
    s0, s1 = out["0"][1], out["1"][1]
    assert s0["attempts"] == s1["attempts"] and s0["rejections"] == s1["rejections"]
    assert abs(s0["t"] - s1["t"]) <= 1e-10 * s0["t"]
    assert np.abs(out["0"][0] - out["1"][0]).max() / 8.6e5 < 1e-10
    for key, v in out["1"][2].items():
        assert v == pytest.approx(out["0"][2][key], rel=1e-10, abs=1e-300)
    sysO = O.System(shape, delta, N=out["0"][3], **kw)
    Mo, so = sysO.run(M, method=method, steps=10, eps_abs=1e-5, eps_rel=1e-5)
    assert s0["attempts"] == so["attempts"] and s0["rejections"] == so["rejections"]
    assert abs(s0["t"] - so["t"]) <= 1e-9 * so["t"]
    assert np.abs(out["0"][0] - Mo).max() / 8.6e5 < 1e-9


@pytest.mark.parametrize("method", ["rkf45", "dp54"])
def test_stepper_split_x_matches_fused(method, monkeypatch):
    # lines too long for three in one workgroup's LDS take a separate inverse
    # x pass (forced here).  Same transforms, but compiled into different
    # kernels, where the compiler may contract multiply-adds differently:
    # agreement to rounding, on a well-conditioned vortex trajectory
    from micromagnum_amd.solver import Stepper
    shape, delta = (24, 20, 3), (3e-9, 3e-9, 2e-9)
    kw = dict(Ms=8.6e5, alpha=0.1, A=1.3e-11, P=0.7, xi=0.05, j=(1e12, 0.0, 0.0))
    M = vortex_field(shape, delta, (36e-9, 30e-9), Ms=8.6e5)
    out = {}
    for split in ("0", "1"):
        monkeypatch.setenv("MM_SPLIT_X", split)
        st = Stepper(shape, delta, precision="fp64", method=method, eps_abs=1e-5, eps_rel=1e-5, **kw)
        st.set_m(M)
        st.run(8)
        out[split] = (st.get_m(), st.stats(), st.energy())
        st.close()
    s0, s1 = out["0"][1], out["1"][1]
    assert s0["attempts"] == s1["attempts"] and s0["rejections"] == s1["rejections"]
    assert abs(s0["t"] - s1["t"]) <= 1e-10 * s0["t"]
    assert np.abs(out["0"][0] - out["1"][0]).max() / 8.6e5 < 1e-10
    for key, v in out["1"][2].items():
        assert v == pytest.approx(out["0"][2][key], rel=1e-10, abs=1e-300)
