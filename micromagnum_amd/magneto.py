"""Drop-in counterparts of the reference's ``magnum.magneto`` functions on the
hot path (src/magneto/bindings/mmm/*.i, math.i:27-66, evolver.i:34-95).

Names, argument order and meaning follow the reference; arrays are
micromagnum_amd.device.DeviceArray objects in the reference's layout (scalar
fields shape (nz, ny, nx), vector fields (3, nz, ny, nx), float64 unless a
``_f32`` entry point is used).  Errors raise RuntimeError like the reference's SWIG layer
(src/magneto/bindings/magneto.i:31-42).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (PADDING_DISABLE, PADDING_ROUND_2, PADDING_ROUND_4, PADDING_ROUND_8,  # noqa: F401
                   PADDING_ROUND_POT, PADDING_SMALL_PRIME_FACTORS)

__all__ = ["exchange", "uniaxial_anisotropy", "cubic_anisotropy", "llge", "fdm_zhangli",
           "scaled_abs_max", "abs_max", "dot_sum", "average", "normalize",
           "GenerateDemagTensor", "demag_tensor_entry", "round_tensor_dimension", "SymmetricMatrixVectorConvolution_FFT",
           "ButcherTableau", "rk_prepare_step", "rk_combine_result",
           "PADDING_DISABLE", "PADDING_ROUND_2", "PADDING_ROUND_4", "PADDING_ROUND_8",
           "PADDING_ROUND_POT", "PADDING_SMALL_PRIME_FACTORS"]


_F64 = np.dtype(np.float64)
_F32 = np.dtype(np.float32)


def _ptr(t, dtype=None):
    from .matty import _Abstract
    if isinstance(t, _Abstract):  # Matrix / VectorMatrix / Field / VectorField
        t = t.device()
    if not hasattr(t, "ptr") or not hasattr(t, "dtype"):
        raise TypeError("expected a micromagnum_amd.device.DeviceArray or a matty Matrix/VectorMatrix")
    if dtype is not None and t.dtype != dtype:
        raise TypeError("expected dtype %s, got %s" % (dtype, t.dtype))
    return t.ptr


def _stream():
    return None


def _f64():
    return _F64


def _n(t):
    from .matty import _Abstract
    if isinstance(t, _Abstract):
        return t.size() * t._comps
    return int(t.numel())


def _rwptr(t, dtype=None):
    """pointer of an argument updated in place"""
    from .matty import _Abstract
    if isinstance(t, _Abstract):
        d = t.device()
        t.mark_device()
        return _ptr(d, dtype)
    return _ptr(t, dtype)


def _wptr(t, dtype=None):
    """pointer of an output argument (a matty object is then marked as
    holding device data)"""
    from .matty import _Abstract
    if isinstance(t, _Abstract):
        d = t.device_out()
        t.mark_device()
        return _ptr(d, dtype)
    return _ptr(t, dtype)


def exchange(*args):
    """exchange(nx,ny,nz,dx,dy,dz,px,py,pz,Ms,A,M,H) -> energy, or
    exchange(Ms, A, M, H) with Fields on a RectangularMesh
    (src/magneto/bindings/mmm/exchange.i:24-39; exchange_cpu.cpp:60-235)."""
    if len(args) == 4:
        Ms, A, M, H = args
        mesh = M.mesh
        nx, ny, nz = mesh.num_nodes
        dx, dy, dz = mesh.delta
        px, py, pz = mesh.pbc
    elif len(args) == 13:
        nx, ny, nz, dx, dy, dz, px, py, pz, Ms, A, M, H = args
    else:
        raise TypeError("exchange() takes 4 or 13 arguments")
    L = _lib.device_lib()
    e = ctypes.c_double(0.0)
    pMs, pA, pM = _ptr(Ms, _f64()), _ptr(A, _f64()), _ptr(M, _f64())
    _lib.check(L.mm_exchange(nx, ny, nz, dx, dy, dz, int(px), int(py), int(pz), pMs, pA, pM, _wptr(H, _f64()),
                             ctypes.byref(e), _stream()))
    return e.value


def exchange_f32(nx, ny, nz, dx, dy, dz, px, py, pz, Ms, A, M, H):
    L = _lib.device_lib()
    _lib.check(L.mm_exchange_f32(nx, ny, nz, dx, dy, dz, int(px), int(py), int(pz), _ptr(Ms, _F32),
                                 _ptr(A, _F32), _ptr(M, _F32), _ptr(H, _F32), _stream()))


def uniaxial_anisotropy(axis, k, Ms, M, H):
    """uniaxial_anisotropy(axis,k,Ms,M,H) -> sum_i k|a x m|^2 (anisotropy.i:24-39)."""
    L = _lib.device_lib()
    e = ctypes.c_double(0.0)
    _lib.check(L.mm_uniaxial_anisotropy(_n(Ms), _ptr(axis, _f64()), _ptr(k, _f64()), _ptr(Ms, _f64()),
                                        _ptr(M, _f64()), _wptr(H, _f64()), ctypes.byref(e), _stream()))
    return e.value


def cubic_anisotropy(axis1, axis2, k, Ms, M, H):
    """cubic_anisotropy(axis1,axis2,k,Ms,M,H) -> sum_i k(a1^2a2^2+...) (anisotropy.i:24-39)."""
    L = _lib.device_lib()
    e = ctypes.c_double(0.0)
    _lib.check(L.mm_cubic_anisotropy(_n(Ms), _ptr(axis1, _f64()), _ptr(axis2, _f64()), _ptr(k, _f64()),
                                     _ptr(Ms, _f64()), _ptr(M, _f64()), _wptr(H, _f64()), ctypes.byref(e),
                                     _stream()))
    return e.value


def llge(f1, f2, M, H, dM):
    """llge(f1,f2,M,H,dM): dM = f1 MxH + f2 Mx(MxH) (llge.i:24-29)."""
    L = _lib.device_lib()
    _lib.check(L.mm_llge(_n(f1), _ptr(f1, _f64()), _ptr(f2, _f64()), _ptr(M, _f64()), _ptr(H, _f64()),
                         _wptr(dM, _f64()), _stream()))


def fdm_zhangli(nx, ny, nz, dx, dy, dz, do_precess, P, xi, Ms, alpha, j, M, dM):
    """fdm_zhangli(...) (zhangli.i:24-32; fdm_zhangli_cpu.cpp:57-200)."""
    L = _lib.device_lib()
    _lib.check(L.mm_fdm_zhangli(nx, ny, nz, dx, dy, dz, int(bool(do_precess)), _ptr(P, _f64()), _ptr(xi, _f64()),
                                _ptr(Ms, _f64()), _ptr(alpha, _f64()), _ptr(j, _f64()), _ptr(M, _f64()),
                                _wptr(dM, _f64()), _stream()))


def _is_uniform(t):
    from .matty import _Abstract
    if isinstance(t, _Abstract):
        return t.isUniform()
    a = t.numpy().reshape(-1)
    return bool(np.all(a == a[0]))


def scaled_abs_max(M, scale, uniform=None):
    """scaled_abs_max(M, scale) (math.i:66; ScaledAbsMax.cpp:52-68)."""
    L = _lib.device_lib()
    if uniform is None:
        uniform = _is_uniform(scale)
    out = ctypes.c_double(0.0)
    _lib.check(L.mm_scaled_abs_max(_n(scale), _ptr(M, _f64()), _ptr(scale, _f64()), int(bool(uniform)),
                                   ctypes.byref(out), _stream()))
    return out.value


def abs_max(M):
    L = _lib.device_lib()
    out = ctypes.c_double(0.0)
    _lib.check(L.mm_abs_max(_n(M) // 3, _ptr(M, _f64()), ctypes.byref(out), _stream()))
    return out.value


def dot_sum(A, B):
    L = _lib.device_lib()
    out = ctypes.c_double(0.0)
    _lib.check(L.mm_dot_sum(_n(A) // 3, _ptr(A, _f64()), _ptr(B, _f64()), ctypes.byref(out), _stream()))
    return out.value


def average(M):
    L = _lib.device_lib()
    out = (ctypes.c_double * 3)()
    _lib.check(L.mm_average(_n(M) // 3, _ptr(M, _f64()), out, _stream()))
    return tuple(out)


def normalize(M, length, mode=0):
    """VectorMatrix.normalize(Matrix len) in place (VectorMatrix.cpp:218-230)."""
    L = _lib.device_lib()
    pl = _ptr(length, _f64())
    pM = _rwptr(M, _f64())
    _lib.check(L.mm_normalize(_n(length), pM, pl, int(mode), _stream()))


def round_tensor_dimension(dim, periodic, padding):
    L = _lib.load()
    out = ctypes.c_int(0)
    _lib.check(L.mm_round_tensor_dimension(int(dim), int(bool(periodic)), int(padding), ctypes.byref(out)))
    return out.value


def GenerateDemagTensor(nx, ny, nz, dx, dy, dz, px=False, py=False, pz=False, repeat=1,
                        padding=PADDING_ROUND_4, cache_dir=None, nthreads=0, device=None, near_on_gpu=False,
                        info=None):
    """Demag tensor field (demag.i:36-42; demag_tensor.cpp:116-186).  Periodic
    axes keep the mesh size and sum `repeat` images plus the far-image
    correction.  Returns a float64 numpy array of shape (ez, ey, ex, 6) whose
    memory order is the reference's Shape(6, ex, ey, ez).

    device="gpu" (open boundaries only): the GPU evaluates the series and
    asymptotic entries in FP64 and the host the finite-difference entries in
    long double (near_on_gpu: those in FP64 on the GPU too); `info`, a dict,
    receives the split and the times (mm_generate_demag_tensor_gpu)."""
    if device == "gpu":
        if px or py or pz:
            raise ValueError("GenerateDemagTensor(device='gpu') takes open boundaries only")
        L = _lib.device_lib()
        ex = (ctypes.c_int * 3)()
        _lib.check(L.mm_generate_demag_tensor_gpu(nx, ny, nz, dx, dy, dz, int(padding), int(bool(near_on_gpu)),
                                                  int(nthreads), None, ex, None))
        N = np.empty((ex[2], ex[1], ex[0], 6), dtype=np.float64)
        inf = (ctypes.c_double * 4)()
        _lib.check(L.mm_generate_demag_tensor_gpu(nx, ny, nz, dx, dy, dz, int(padding), int(bool(near_on_gpu)),
                                                  int(nthreads), N.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  ex, inf))
        if info is not None:
            info.update(gpu_entries=int(inf[0]), host_entries=int(inf[1]), gpu_s=inf[2], host_s=inf[3])
        return N
    L = _lib.load()
    ex = (ctypes.c_int * 3)()
    cd = cache_dir.encode() if cache_dir else None
    _lib.check(L.mm_generate_demag_tensor(nx, ny, nz, dx, dy, dz, int(px), int(py), int(pz), int(repeat),
                                          int(padding), cd, int(nthreads), None, ex))
    N = np.empty((ex[2], ex[1], ex[0], 6), dtype=np.float64)
    _lib.check(L.mm_generate_demag_tensor(nx, ny, nz, dx, dy, dz, int(px), int(py), int(pz), int(repeat),
                                          int(padding), cd, int(nthreads),
                                          N.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ex))
    return N


def demag_tensor_entry(i, j, k, dx, dy, dz):
    """The six components (xx, xy, xz, yy, yz, zz) of the open-boundary
    tensor at (i, j, k) of the padded field GenerateDemagTensor returns,
    computed on the host in long double (mm_demag_tensor_entry)."""
    v = (ctypes.c_double * 6)()
    _lib.check(_lib.load().mm_demag_tensor_entry(int(i), int(j), int(k), dx, dy, dz, v))
    return np.array(v[:])


class SymmetricMatrixVectorConvolution_FFT:
    """SymmetricMatrixVectorConvolution_FFT(N, nx, ny, nz) (math.i:27-33)."""

    def __init__(self, N, nx, ny, nz, precision="fp64"):
        L = _lib.device_lib()
        N = np.ascontiguousarray(N, dtype=np.float64)
        ez, ey, ex, six = N.shape
        assert six == 6
        self.precision = _lib.MM_FP64 if precision == "fp64" else _lib.MM_FP32
        self.shape = (nx, ny, nz)
        h = ctypes.c_void_p()
        _lib.check(L.mm_conv_create(N.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nx, ny, nz, ex, ey, ez,
                                    self.precision, ctypes.byref(h)))
        self._h = h

    @property
    def fft_size(self):
        out = (ctypes.c_int * 3)()
        _lib.check(_lib.load().mm_conv_fft_size(self._h, out))
        return tuple(out)

    def execute(self, M, H):
        L = _lib.device_lib()
        if self.precision == _lib.MM_FP64:
            pM = _ptr(M, _F64)
            _lib.check(L.mm_conv_execute(self._h, pM, _wptr(H, _F64), _stream()))
        else:
            _lib.check(L.mm_conv_execute_f32(self._h, _ptr(M, _F32), _ptr(H, _F32), _stream()))

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib._LIB is not None:
                _lib._LIB.mm_conv_destroy(self._h)
                self._h = None
        except Exception:
            pass


class SymmetricMatrixVectorConvolution_Simple:
    """SymmetricMatrixVectorConvolution_Simple(N, nx, ny, nz) (math.i:35-41):
    the direct-sum convolution, FP64, O(n^2)."""

    def __init__(self, N, nx, ny, nz):
        L = _lib.device_lib()
        N = np.ascontiguousarray(N, dtype=np.float64)
        ez, ey, ex, six = N.shape
        assert six == 6
        self.shape = (nx, ny, nz)
        h = ctypes.c_void_p()
        _lib.check(L.mm_conv_simple_create(N.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nx, ny, nz, ex, ey,
                                           ez, ctypes.byref(h)))
        self._h = h

    def execute(self, M, H):
        _lib.check(_lib.device_lib().mm_conv_simple_execute(self._h, _ptr(M, _F64), _wptr(H, _F64), _stream()))

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib._LIB is not None:
                _lib._LIB.mm_conv_simple_destroy(self._h)
                self._h = None
        except Exception:
            pass


class ButcherTableau:
    """Butcher tableau of a named method (runge_kutta.h:31-72, tableaus.py:54-139)."""

    def __init__(self, method):
        self._t = _lib.Tableau()
        _lib.check(_lib.load().mm_tableau(_lib.METHODS.get(method, method), ctypes.byref(self._t)))

    def getNumSteps(self):
        return self._t.num_steps

    def getA(self, i):
        return self._t.a[i]

    def getB(self, i, j):
        return self._t.b[i][j]

    def getC(self, i):
        return self._t.c[i]

    def getEC(self, i):
        return self._t.ec[i]

    @property
    def order(self):
        return self._t.order

    @property
    def fsal(self):
        return bool(self._t.fsal)


def _kptrs(ks):
    arr = (ctypes.c_void_p * 8)()
    for q, k in enumerate(ks):
        arr[q] = None if k is None else _ptr(k, _F64).value
    return arr


def rk_prepare_step(step, h, tab, k, y, ytmp):
    """rk_prepare_step(step, h, tab, k0..k5, y, ytmp) (evolver.i:34-95); k is a list."""
    L = _lib.device_lib()
    ks, py = _kptrs(k), _ptr(y, _f64())
    _lib.check(L.mm_rk_prepare_step(int(step), float(h), ctypes.byref(tab._t), _n(y) // 3, ks, py,
                                    _wptr(ytmp, _f64()), _stream()))


def rk_combine_result(h, tab, k, y, y_err):
    """rk_combine_result(h, tab, k0..k5, y, y_err) (evolver.i:34-95); k is a list."""
    L = _lib.device_lib()
    ks = _kptrs(k)
    _lib.check(L.mm_rk_combine_result(float(h), ctypes.byref(tab._t), _n(y) // 3, ks, _rwptr(y, _f64()),
                                      _wptr(y_err, _f64()), _stream()))
