"""magneto_amd -- the hot-path functions of the reference's ``magneto`` module
on libmmhip.so, for the reference's own objects.

This is the reference-side binding.  A maintainer copies this one file to
``src/magnum/magneto_amd.py`` and tries it first where the reference picks its
native module (src/magnum/magneto.py:19-21).  It imports nothing from this
package: the matrices it takes are the reference's SWIG proxies (Matrix,
VectorMatrix, Field, VectorField, ButcherTableau), and data crosses through
accessors the reference binds (src/magneto/matrix/matty.inc.i):

  VectorMatrix.toByteArray(c)     component c, linear index (x fastest)  :276-288
  Matrix.toByteArray()            linear index (x fastest)               :175-187
  VectorMatrix.set(i, (x, y, z))  one cell                                :223-229
  Matrix.set(i, v)                one cell                                :139-145
  getShape().getDim(k)            dimensions                              :92

The library's arrays use the same order (SoA planes, x fastest,
src/magneto/matrix/Shape.cpp:64-75), so no strided view is involved.  (The
reference's ``to_numpy`` gives an [x, y, z, c]-indexed view; its SoA form is
``a.transpose(3, 2, 1, 0)`` -- see ``soa_from_to_numpy``.)  The reference has
no bulk setter (its ``from_numpy`` raises NotImplementedError,
matty.inc.i:356-380); results are written back with ``fromByteArray(data,
component)`` / ``fromByteArray(data)`` where the maintainer adds those two
%extend methods (the counterparts of ``toByteArray``), else cell by cell
with ``set``.

Errors raise RuntimeError, as the reference's SWIG layer does
(src/magneto/bindings/magneto.i:31-42).
"""
import ctypes
import os

import numpy as np

try:  # in the reference tree: Matrix, VectorMatrix, Shape, ... stay the CPU module's
    from .magneto_cpu import *  # noqa: F401,F403
except (ImportError, ValueError, SystemError):
    pass

_vp, _dp, _ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
_d, _i = ctypes.c_double, ctypes.c_int
_L = None


def _lib():
    """libmmhip.so, loaded and initialised on first use (MM_LIB: its path,
    MM_DEVICE: the HIP device)."""
    global _L
    if _L is None:
        L = ctypes.CDLL(os.environ.get("MM_LIB", "libmmhip.so"))
        L.mm_last_error.restype = ctypes.c_char_p
        if L.mm_init(int(os.environ.get("MM_DEVICE", "0"))) != 0:
            raise ImportError(L.mm_last_error().decode())
        _L = L
    return _L


def _check(rc):
    if rc != 0:
        raise RuntimeError(_lib().mm_last_error().decode())


# ------------------------------------------------------------ data transfer
def dims(m):
    """Dimensions of a reference matrix (Shape.getDim over its rank)."""
    sh = m.getShape()
    return tuple(sh.getDim(k) for k in range(sh.getRank()))


def vector_planes(vm):
    """VectorMatrix -> float64 array of 3 SoA planes (x fastest): the
    library's vector layout."""
    return np.concatenate([np.frombuffer(vm.toByteArray(c), dtype=np.float64) for c in range(3)])


def matrix_values(m):
    """Matrix -> float64 array in linear-index order (x fastest; a Shape(6,
    ex, ey, ez) tensor comes out component fastest)."""
    return np.frombuffer(m.toByteArray(), dtype=np.float64).copy()


def soa_from_to_numpy(a):
    """The reference's to_numpy() view of a VectorMatrix, indexed [x, y, z, c]
    (matty.inc.i:342-354) -> SoA planes (3, nz, ny, nx), x fastest."""
    return np.ascontiguousarray(np.asarray(a).transpose(3, 2, 1, 0))


def store_vector(vm, planes):
    """Write 3 SoA planes back into a VectorMatrix."""
    planes = np.ascontiguousarray(planes, dtype=np.float64).reshape(3, -1)
    if hasattr(vm, "fromByteArray"):
        for c in range(3):
            vm.fromByteArray(planes[c].tobytes(), c)
        return
    for i in range(planes.shape[1]):
        vm.set(i, (float(planes[0, i]), float(planes[1, i]), float(planes[2, i])))


def store_matrix(m, values):
    values = np.ascontiguousarray(values, dtype=np.float64).ravel()
    if hasattr(m, "fromByteArray"):
        m.fromByteArray(values.tobytes())
        return
    for i in range(values.size):
        m.set(i, float(values[i]))


class _Dev(object):
    """Device copy of a float64 host array."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = _vp()
        n = max(self.host.nbytes, 8)
        _check(_lib().mm_device_alloc(ctypes.c_size_t(n), ctypes.byref(self.ptr)))
        if self.host.nbytes:
            _check(_lib().mm_memcpy_htod(self.ptr, self.host.ctypes.data_as(_vp),
                                         ctypes.c_size_t(self.host.nbytes), None))

    def back(self):
        out = np.empty_like(self.host)
        _check(_lib().mm_memcpy_dtoh(out.ctypes.data_as(_vp), self.ptr, ctypes.c_size_t(out.nbytes), None))
        return out

    def __del__(self):
        if _L is not None and self.ptr:
            _L.mm_device_free(self.ptr)


def _vec(vm):
    return _Dev(vector_planes(vm))


def _sca(m):
    return _Dev(matrix_values(m))


# ------------------------------------------------------------ field terms
def exchange(*args):
    """exchange(nx, ny, nz, dx, dy, dz, px, py, pz, Ms, A, M, H) or
    exchange(Ms, A, M, H) on Fields (bindings/mmm/exchange.i:24-39).
    Returns the exchange energy -mu0/2 dV sum m.H; H is overwritten."""
    if len(args) == 4:
        Ms, A, M, H = args
        mesh = M.getMesh()
        nx, ny, nz = mesh.getNumNodes()
        dx, dy, dz = mesh.getDelta()
        pbc = mesh.getPeriodicBC()[0]
        args = (nx, ny, nz, dx, dy, dz, "x" in pbc, "y" in pbc, "z" in pbc, Ms, A, M, H)
    nx, ny, nz, dx, dy, dz, px, py, pz, Ms, A, M, H = args
    Ms_, A_, M_, H_ = _sca(Ms), _sca(A), _vec(M), _vec(H)
    e = _d()
    _check(_lib().mm_exchange(nx, ny, nz, _d(dx), _d(dy), _d(dz), int(bool(px)), int(bool(py)),
                              int(bool(pz)), Ms_.ptr, A_.ptr, M_.ptr, H_.ptr, ctypes.byref(e), None))
    store_vector(H, H_.back())
    return e.value


def uniaxial_anisotropy(axis, k, Ms, M, H):
    """bindings/mmm/anisotropy.i:24-30; returns sum_i k |a x m|^2."""
    n = M.size()
    ax_, k_, Ms_, M_, H_ = _vec(axis), _sca(k), _sca(Ms), _vec(M), _vec(H)
    e = _d()
    _check(_lib().mm_uniaxial_anisotropy(n, ax_.ptr, k_.ptr, Ms_.ptr, M_.ptr, H_.ptr, ctypes.byref(e), None))
    store_vector(H, H_.back())
    return e.value


def cubic_anisotropy(axis1, axis2, k, Ms, M, H):
    """bindings/mmm/anisotropy.i:32-39."""
    n = M.size()
    a1_, a2_, k_, Ms_, M_, H_ = _vec(axis1), _vec(axis2), _sca(k), _sca(Ms), _vec(M), _vec(H)
    e = _d()
    _check(_lib().mm_cubic_anisotropy(n, a1_.ptr, a2_.ptr, k_.ptr, Ms_.ptr, M_.ptr, H_.ptr,
                                      ctypes.byref(e), None))
    store_vector(H, H_.back())
    return e.value


def llge(f1, f2, M, H, dM):
    """bindings/mmm/llge.i:24-29: dM = f1 M x H + f2 M x (M x H)."""
    f1_, f2_, M_, H_, dM_ = _sca(f1), _sca(f2), _vec(M), _vec(H), _vec(dM)
    _check(_lib().mm_llge(M.size(), f1_.ptr, f2_.ptr, M_.ptr, H_.ptr, dM_.ptr, None))
    store_vector(dM, dM_.back())


def fdm_zhangli(nx, ny, nz, dx, dy, dz, do_precess, P, xi, Ms, alpha, j, M, dM):
    """bindings/mmm/zhangli.i:24-32."""
    P_, xi_, Ms_, al_, j_, M_, dM_ = (_sca(P), _sca(xi), _sca(Ms), _sca(alpha), _vec(j), _vec(M), _vec(dM))
    _check(_lib().mm_fdm_zhangli(nx, ny, nz, _d(dx), _d(dy), _d(dz), int(bool(do_precess)), P_.ptr,
                                 xi_.ptr, Ms_.ptr, al_.ptr, j_.ptr, M_.ptr, dM_.ptr, None))
    store_vector(dM, dM_.back())


def scaled_abs_max(M, scale):
    """bindings/math.i:66: max_i |M_i| / scale_i (ScaledAbsMax.cpp:33-68)."""
    s = matrix_values(scale)
    uniform = bool(s.size == 0 or np.all(s == s[0]))
    M_, s_ = _vec(M), _Dev(s)
    out = _d()
    _check(_lib().mm_scaled_abs_max(M.size(), M_.ptr, s_.ptr, int(uniform), ctypes.byref(out), None))
    return out.value


# ------------------------------------------------------------ demag
PADDING_DISABLE, PADDING_ROUND_2, PADDING_ROUND_4, PADDING_ROUND_8 = 0, 1, 2, 3
PADDING_ROUND_POT, PADDING_SMALL_PRIME_FACTORS = 4, 5


def GenerateDemagTensor(dim_x, dim_y, dim_z, delta_x, delta_y, delta_z, periodic_x, periodic_y,
                        periodic_z, periodic_repeat, padding, cache_dir):
    """bindings/mmm/demag.i:36-42 (demag_tensor.cpp:116-186): the reference's
    Matrix(Shape(6, ex, ey, ez)), sign-flipped so that H = N (*) M."""
    L = _lib()
    ex = (_i * 3)()
    cd = cache_dir.encode() if cache_dir else None
    args = (dim_x, dim_y, dim_z, _d(delta_x), _d(delta_y), _d(delta_z), int(bool(periodic_x)),
            int(bool(periodic_y)), int(bool(periodic_z)), int(periodic_repeat), int(padding), cd, 0)
    _check(L.mm_generate_demag_tensor(*args, None, ex))
    N = np.empty(6 * ex[0] * ex[1] * ex[2])
    _check(L.mm_generate_demag_tensor(*args, N.ctypes.data_as(_dp), ex))
    m = Matrix(Shape(6, ex[0], ex[1], ex[2]))  # noqa: F405 (the reference's classes)
    store_matrix(m, N)
    return m


class SymmetricMatrixVectorConvolution_FFT(object):
    """bindings/math.i:27-33.  N: the reference's Shape(6, ex, ey, ez) tensor
    (GenerateDemagTensor), component fastest."""

    _create, _execute = "mm_conv_create", "mm_conv_execute"

    def __init__(self, N, dim_x, dim_y, dim_z):
        _, ex, ey, ez = dims(N)
        self._N = matrix_values(N)
        self._h = _vp()
        self._make(dim_x, dim_y, dim_z, ex, ey, ez)

    def _make(self, nx, ny, nz, ex, ey, ez):
        _check(_lib().mm_conv_create(self._N.ctypes.data_as(_dp), nx, ny, nz, ex, ey, ez, 0,
                                     ctypes.byref(self._h)))

    def execute(self, rhs, res):
        M_, H_ = _vec(rhs), _vec(res)
        _check(getattr(_lib(), self._execute)(self._h, M_.ptr, H_.ptr, None))
        store_vector(res, H_.back())

    def __del__(self):
        if _L is not None and self._h:
            _L.mm_conv_destroy(self._h)


class SymmetricMatrixVectorConvolution_Simple(SymmetricMatrixVectorConvolution_FFT):
    """bindings/math.i:35-41: the direct sum."""

    _execute = "mm_conv_simple_execute"

    def _make(self, nx, ny, nz, ex, ey, ez):
        _check(_lib().mm_conv_simple_create(self._N.ctypes.data_as(_dp), nx, ny, nz, ex, ey, ez,
                                            ctypes.byref(self._h)))

    def __del__(self):
        if _L is not None and self._h:
            _L.mm_conv_simple_destroy(self._h)


# ------------------------------------------------------------ Runge-Kutta algebra
class _Tab(ctypes.Structure):
    _fields_ = [("num_steps", _i), ("order", _i), ("fsal", _i),
                ("a", _d * 8), ("b", (_d * 8) * 8), ("c", _d * 8), ("ec", _d * 8)]


def _tableau(tab):
    """the reference's ButcherTableau (bindings/evolver.i:34-60) as mm_tableau_t"""
    t = _Tab()
    n = tab.getNumSteps()
    t.num_steps = n
    for i in range(n):
        t.a[i] = tab.getA(i)
        t.c[i] = tab.getC(i)
        t.ec[i] = tab.getEC(i)
        for j in range(i):
            t.b[i][j] = tab.getB(i, j)
    return t


def rk_prepare_step(step, h, tab, k0, k1, k2, k3, k4, k5, y, ytmp):
    """bindings/evolver.i:62-75: ytmp = y + h sum_{j<step} b[step][j] k_j."""
    ks = [_vec(k) for k in (k0, k1, k2, k3, k4, k5)[:step]]
    y_, yt_ = _vec(y), _vec(ytmp)
    arr = (_vp * 6)(*([k.ptr for k in ks] + [None] * (6 - len(ks))))
    t = _tableau(tab)
    _check(_lib().mm_rk_prepare_step(step, _d(h), ctypes.byref(t), y.size(), arr, y_.ptr, yt_.ptr, None))
    store_vector(ytmp, yt_.back())


def rk_combine_result(h, tab, k0, k1, k2, k3, k4, k5, y, y_error):
    """bindings/evolver.i:77-90: y += h sum c_j k_j; y_error = h sum ec_j k_j."""
    ks = [_vec(k) for k in (k0, k1, k2, k3, k4, k5)]
    y_, e_ = _vec(y), _vec(y_error)
    arr = (_vp * 6)(*[k.ptr for k in ks])
    t = _tableau(tab)
    _check(_lib().mm_rk_combine_result(_d(h), ctypes.byref(t), y.size(), arr, y_.ptr, e_.ptr, None))
    store_vector(y, y_.back())
    store_vector(y_error, e_.back())


# ------------------------------------------------------------ fused evolver
class SystemDesc(ctypes.Structure):
    """mm_system_desc_t (include/mm_hotpath.h)."""
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("dx", _d), ("dy", _d), ("dz", _d),
                ("precision", _i), ("method", _i), ("eps_abs", _d), ("eps_rel", _d), ("h0", _d),
                ("normalize_mode", _i), ("do_precess", _i),
                ("Ms_host", _dp), ("alpha_host", _dp), ("A_host", _dp),
                ("k_uni_host", _dp), ("k_cub_host", _dp), ("axis1_host", _dp), ("axis2_host", _dp),
                ("P_host", _dp), ("xi_host", _dp), ("j_host", _dp),
                ("use_zeeman", _i), ("h_ext", _d * 3),
                ("N_host", _dp), ("ex", _i), ("ey", _i), ("ez", _i),
                ("px", _i), ("py", _i), ("pz", _i),
                ("use_ac", _i), ("h_amp", _d * 3), ("h_freq", _d * 3), ("h_phase", _d * 3)]


_METHODS = {"rkf45": 0, "cc45": 1, "dp54": 2, "rk23": 3, "rk4": 4, "euler": 5}


class AmdRungeKutta(object):
    """RungeKutta.evolve (src/magnum/evolver/runge_kutta.py:47-87) with the
    step controller of micro_magnetics_stepsize_controller.py:35-73 and
    finish_step's renormalisation (micro_magnetics.py:49-50) on the fused
    device stepper.  ``desc`` is a filled SystemDesc (its host arrays must
    outlive the constructor call only).

    evolve(state, t_max) makes one accepted step and clips it to t_max as the
    reference does (runge_kutta.py:74-78).  M is copied to the device only
    when state.y changed since the last step, so the FSAL stage of dp54/rk23
    carries over between steps as in the reference (runge_kutta.py:97-101).
    As there, the next trial step belongs to the state: a state the evolver
    has not stepped before starts from state.h (runge_kutta.py:55-58)."""

    def __init__(self, desc):
        self.st = _vp()
        _check(_lib().mm_stepper_create(ctypes.byref(desc), ctypes.byref(self.st)))
        self._last = None
        self._state = None
        self._t = None

    def evolve(self, state, t_max):
        L = _lib()
        y = vector_planes(state.y)
        fresh = self._state is not state
        if fresh or self._last is None or not np.array_equal(y, self._last):
            _check(L.mm_stepper_set_m(self.st, y.ctypes.data_as(_dp)))
        if fresh or state.t != self._t:
            h = state.h if fresh and state.h > 0 else 0.0
            _check(L.mm_stepper_set_time(self.st, _d(state.t), _d(h)))
        self._state = state
        _check(L.mm_stepper_run_until(self.st, _d(t_max), ctypes.c_longlong(1)))
        out = np.empty_like(y)
        _check(L.mm_stepper_get_m(self.st, out.ctypes.data_as(_dp)))
        store_vector(state.y, out)
        self._last = out
        stats = (_d * 5)()
        _check(L.mm_stepper_stats(self.st, stats))
        state.t, state.h = stats[0], stats[1]
        self._t = state.t
        state.step += 1
        state.substep = 0
        state.flush_cache()
        return state

    def __del__(self):
        if _L is not None and self.st:
            _L.mm_stepper_destroy(self.st)
