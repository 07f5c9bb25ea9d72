"""ctypes front end of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py``; never by the product package.

Vector fields are numpy float64 arrays of shape (3, nz, ny, nx) (SoA, x
fastest, the reference's VectorMatrix layout); scalar fields (nz, ny, nx).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

RKF45, CC45, DP54, RK23, RK4, EULER = 0, 1, 2, 3, 4, 5
METHODS = {"rkf45": RKF45, "cc45": CC45, "dp54": DP54, "rk23": RK23, "rk4": RK4, "euler": EULER}

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liborc.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_mu0.restype = ctypes.c_double
        L.orc_exchange.restype = ctypes.c_double
        L.orc_uniaxial_anisotropy.restype = ctypes.c_double
        L.orc_cubic_anisotropy.restype = ctypes.c_double
        L.orc_scaled_abs_max.restype = ctypes.c_double
        L.orc_abs_max.restype = ctypes.c_double
        L.orc_dot_sum.restype = ctypes.c_double
        L.orc_conv_create.restype = ctypes.c_void_p
        L.orc_conv_create.argtypes = [_dp] + [ctypes.c_int] * 6
        L.orc_conv_execute.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_conv_destroy.argtypes = [ctypes.c_void_p]
        L.orc_conv_get_kernel.argtypes = [ctypes.c_void_p, _dp, _dp]
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "expect contiguous float64"
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mu0():
    return lib().orc_mu0()


def set_threads(n):
    lib().orc_set_threads(int(n))


def fft(x, sign=-1):
    """Unnormalised complex DFT of a 1-D complex array (sign -1 forward)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    buf = x.view(np.float64)
    lib().orc_fft(_p(buf), ctypes.c_int(x.size), ctypes.c_int(sign))
    return x


class Conv:
    """SymmetricMatrixVectorConvolution_FFT restated on the CPU."""

    def __init__(self, N, nx, ny, nz):
        N = _c(N)
        self.shape = (nx, ny, nz)
        Pz, Py, Px = N.shape[0], N.shape[1], N.shape[2]
        assert N.shape[3] == 6
        self.P = (Px, Py, Pz)
        self._N = N
        self.h = lib().orc_conv_create(_p(N), nx, ny, nz, Px, Py, Pz)

    def execute(self, M):
        M = _c(M)
        H = np.empty_like(M)
        lib().orc_conv_execute(self.h, _p(M), _p(H))
        return H

    def kernel(self):
        Px, Py, Pz = self.P
        Hx = Px // 2 + 1
        re = np.empty((Pz, Py, Hx, 6))
        im = np.empty((Pz, Py, Hx, 6))
        lib().orc_conv_get_kernel(self.h, _p(re), _p(im))
        return re, im

    def __del__(self):
        try:
            if self.h:
                lib().orc_conv_destroy(self.h)
        except Exception:
            pass


def conv_simple(N, M):
    """Direct-sum convolution (SymmetricMatrixVectorConvolution_Simple)."""
    N = _c(N)
    M = _c(M)
    _, nz, ny, nx = M.shape
    Pz, Py, Px = N.shape[:3]
    H = np.empty_like(M)
    lib().orc_conv_simple(_p(N), nx, ny, nz, Px, Py, Pz, _p(M), _p(H))
    return H


def exchange(delta, Ms, A, M, pbc=(False, False, False)):
    Ms, A, M = _c(Ms), _c(A), _c(M)
    _, nz, ny, nx = M.shape
    H = np.empty_like(M)
    dx, dy, dz = (ctypes.c_double(v) for v in delta)
    E = lib().orc_exchange(nx, ny, nz, dx, dy, dz, int(pbc[0]), int(pbc[1]), int(pbc[2]),
                           _p(Ms), _p(A), _p(M), _p(H))
    return H, E


def uniaxial_anisotropy(axis, k, Ms, M):
    axis, k, Ms, M = _c(axis), _c(k), _c(Ms), _c(M)
    H = np.empty_like(M)
    E = lib().orc_uniaxial_anisotropy(Ms.size, _p(axis), _p(k), _p(Ms), _p(M), _p(H))
    return H, E


def cubic_anisotropy(axis1, axis2, k, Ms, M):
    axis1, axis2, k, Ms, M = _c(axis1), _c(axis2), _c(k), _c(Ms), _c(M)
    H = np.empty_like(M)
    E = lib().orc_cubic_anisotropy(Ms.size, _p(axis1), _p(axis2), _p(k), _p(Ms), _p(M), _p(H))
    return H, E


def llge(f1, f2, M, H):
    f1, f2, M, H = _c(f1), _c(f2), _c(M), _c(H)
    dM = np.empty_like(M)
    lib().orc_llge(f1.size, _p(f1), _p(f2), _p(M), _p(H), _p(dM))
    return dM


def zhangli(delta, do_precess, P, xi, Ms, alpha, j, M):
    P, xi, Ms, alpha, j, M = (_c(a) for a in (P, xi, Ms, alpha, j, M))
    _, nz, ny, nx = M.shape
    dM = np.empty_like(M)
    dx, dy, dz = (ctypes.c_double(v) for v in delta)
    lib().orc_zhangli(nx, ny, nz, dx, dy, dz, int(bool(do_precess)), _p(P), _p(xi), _p(Ms),
                      _p(alpha), _p(j), _p(M), _p(dM))
    return dM


def scaled_abs_max(M, s, uniform=None):
    M, s = _c(M), _c(s)
    if uniform is None:
        uniform = bool(np.all(s == s.flat[0]))
    return lib().orc_scaled_abs_max(s.size, _p(M), _p(s), int(uniform))


def controller(eps_abs, eps_rel, h, order, max_y_err, max_dydt):
    """MicroMagneticsStepSizeController.adjust_stepsize on the two maxima:
    returns (accept, h_next)."""
    L = lib()
    d = ctypes.c_double
    hn = d(0.0)
    acc = L.orc_controller(d(eps_abs), d(eps_rel), d(h), int(order), d(max_y_err), d(max_dydt),
                           ctypes.byref(hn))
    return bool(acc), hn.value


def normalize(M, length, mode=0):
    M = _c(M).copy()
    length = _c(np.broadcast_to(length, M.shape[1:]))
    lib().orc_normalize(length.size, _p(M), _p(length), int(mode))
    return M


class _Tableau(ctypes.Structure):
    _fields_ = [("num_steps", ctypes.c_int), ("order", ctypes.c_int), ("fsal", ctypes.c_int),
                ("a", ctypes.c_double * 7), ("b", (ctypes.c_double * 7) * 7),
                ("c", ctypes.c_double * 7), ("ec", ctypes.c_double * 7)]


def tableau(method):
    t = _Tableau()
    lib().orc_tableau_get(METHODS.get(method, method), ctypes.byref(t))
    n = t.num_steps
    return dict(num_steps=n, order=t.order, fsal=t.fsal,
                a=np.array(t.a[:n]), b=np.array([[t.b[i][j] for j in range(n)] for i in range(n)]),
                c=np.array(t.c[:n]), ec=np.array(t.ec[:n]))


class _System(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("px", ctypes.c_int), ("py", ctypes.c_int), ("pz", ctypes.c_int),
                ("Ms", _dp), ("A", _dp), ("alpha", _dp), ("Ms_uniform", ctypes.c_int),
                ("conv", ctypes.c_void_p), ("use_exchange", ctypes.c_int),
                ("k_uni", _dp), ("k_cub", _dp), ("axis1", _dp), ("axis2", _dp),
                ("use_zeeman", ctypes.c_int), ("h_ext", ctypes.c_double * 3),
                ("P", _dp), ("xi", _dp), ("j", _dp), ("do_precess", ctypes.c_int),
                ("normalize_mode", ctypes.c_int), ("use_ac", ctypes.c_int),
                ("h_amp", ctypes.c_double * 3), ("h_freq", ctypes.c_double * 3),
                ("h_phase", ctypes.c_double * 3),
                ("h_fn", ctypes.CFUNCTYPE(None, ctypes.c_double, _dp, ctypes.c_void_p)),
                ("h_user", ctypes.c_void_p)]


class System:
    """Micromagnetic system in the module order stray, exchange, anisotropy,
    Zeeman, spin torque (the order the tests pass to create_solver)."""

    def __init__(self, shape, delta, Ms, alpha, A=None, N=None, k_uni=None, k_cub=None,
                 axis1=None, axis2=None, h_ext=None, P=None, xi=None, j=None, do_precess=True,
                 normalize_mode=0, pbc=(False, False, False), h_ext_amp=None, h_ext_freq=None,
                 h_ext_phase=None, h_ext_fn=None):
        """h_ext is the Zeeman offset; h_ext_amp/freq/phase (3-tuples) make it
        the reference's AlternatingField, h_ext + amp sin(t freq + phase);
        h_ext_fn(t) -> 3-vector replaces them (AlternatingField with a user
        function)."""
        nx, ny, nz = shape
        self._keep = []

        def field(a, comps=1):
            if a is None:
                return None
            a = np.asarray(a, dtype=np.float64)
            if comps > 1 and a.ndim == 1:
                a = a.reshape(comps, 1, 1, 1)
            a = _c(np.broadcast_to(a, ((comps,) if comps > 1 else ()) + (nz, ny, nx)))
            self._keep.append(a)
            return a

        s = _System()
        s.nx, s.ny, s.nz = nx, ny, nz
        s.dx, s.dy, s.dz = delta
        s.px, s.py, s.pz = (int(b) for b in pbc)
        Ms_a = field(Ms)
        s.Ms = _p(Ms_a)
        s.Ms_uniform = int(np.all(Ms_a == Ms_a.flat[0]))
        s.alpha = _p(field(alpha))
        s.A = _p(field(A if A is not None else 0.0))
        s.use_exchange = int(A is not None)
        self.conv = None
        if N is not None:
            self.conv = Conv(N, nx, ny, nz)
            s.conv = self.conv.h
        s.k_uni = _p(field(k_uni))
        s.k_cub = _p(field(k_cub))
        s.axis1 = _p(field(axis1, 3))
        s.axis2 = _p(field(axis2, 3))
        if h_ext is not None or h_ext_amp is not None:
            s.use_zeeman = 1
            for c in range(3):
                s.h_ext[c] = float(h_ext[c]) if h_ext is not None else 0.0
        if h_ext_amp is not None:
            s.use_ac = 1
            for c in range(3):
                s.h_amp[c] = float(h_ext_amp[c])
                s.h_freq[c] = float(h_ext_freq[c]) if h_ext_freq is not None else 0.0
                s.h_phase[c] = float(h_ext_phase[c]) if h_ext_phase is not None else 0.0
        if h_ext_fn is not None:
            s.use_zeeman = 1

            def cb(t, h, user):
                v = h_ext_fn(t)
                h[0], h[1], h[2] = float(v[0]), float(v[1]), float(v[2])

            self._fn_cb = _System._fields_[-2][1](cb)
            s.h_fn = self._fn_cb
        s.P = _p(field(P))
        s.xi = _p(field(xi))
        s.j = _p(field(j, 3))
        s.do_precess = int(bool(do_precess))
        s.normalize_mode = int(normalize_mode)
        self.s = s
        self.n = nx * ny * nz
        self.shape = shape

    def llge_factors(self):
        f1 = np.empty(self.n)
        f2 = np.empty(self.n)
        lib().orc_llge_factors(ctypes.byref(self.s), _p(f1), _p(f2))
        return f1, f2

    def h_tot(self, M, t=0.0):
        M = _c(M)
        H = np.empty_like(M)
        lib().orc_h_tot_at(ctypes.byref(self.s), _p(M), ctypes.c_double(t), _p(H))
        return H

    def dmdt(self, M, t=0.0):
        M = _c(M)
        f1, f2 = self.llge_factors()
        dM = np.empty_like(M)
        H = np.empty_like(M)
        lib().orc_dmdt_at(ctypes.byref(self.s), _p(f1), _p(f2), _p(M), ctypes.c_double(t), _p(dM), _p(H))
        return dM, H

    def run(self, M, method="rkf45", steps=1, eps_abs=1e-4, eps_rel=1e-4, h0=1e-13,
            record_avg=False):
        M = _c(M).copy()
        stats = np.zeros(5)
        avg = np.zeros((steps, 3)) if record_avg else None
        lib().orc_run(ctypes.byref(self.s), METHODS[method], ctypes.c_double(eps_abs),
                      ctypes.c_double(eps_rel), ctypes.c_double(h0), steps, _p(M), _p(stats),
                      _p(avg) if avg is not None else None)
        out = dict(t=stats[0], h=stats[1], h_next=stats[2], attempts=int(stats[3]),
                   rejections=int(stats[4]))
        if avg is not None:
            out["avg"] = avg
        return M, out

    ENERGY_TERMS = ("E_stray", "E_exch", "E_aniso", "E_ext", "E_tot")

    def energy(self, M, t=0.0):
        """E_stray, E_exch, E_aniso, E_ext, E_tot at M and time t (joule)."""
        M = _c(M)
        e = np.zeros(5)
        lib().orc_energy_at(ctypes.byref(self.s), _p(M), ctypes.c_double(t), _p(e))
        return dict(zip(self.ENERGY_TERMS, e.tolist()))

    def relax(self, M, method="rkf45", threshold=1e7, criterion="dmdt", check_every=1,
              max_steps=100000, eps_abs=1e-4, eps_rel=1e-4, h0=1e-13):
        """Run until relaxed (criterion "dmdt": scaled_abs_max(dM/dt, Ms) <
        threshold; "deg_per_ns": the reference's Relaxed condition)."""
        M = _c(M).copy()
        stats = np.zeros(5)
        info = np.zeros(3)
        crit = {"dmdt": 1, "deg_per_ns": 2}[criterion]
        L = lib()
        L.orc_relax.restype = ctypes.c_long
        L.orc_relax(ctypes.byref(self.s), METHODS[method], ctypes.c_double(eps_abs),
                    ctypes.c_double(eps_rel), ctypes.c_double(h0), crit, ctypes.c_double(threshold),
                    int(check_every), ctypes.c_long(int(max_steps)), _p(M), _p(stats), _p(info))
        out = dict(t=stats[0], h=stats[1], h_next=stats[2], attempts=int(stats[3]),
                   rejections=int(stats[4]), steps=int(info[0]), relaxed=bool(info[1]),
                   value=float(info[2]))
        return M, out

    def run_to_time(self, M, t_stop, method="rkf45", max_steps=1000000, eps_abs=1e-4, eps_rel=1e-4, h0=1e-13):
        """Accepted steps until the first one that ends at t >= t_stop, none
        clipped (the reference's solve(Time(t_stop)))."""
        M = _c(M).copy()
        stats = np.zeros(5)
        L = lib()
        L.orc_run_to_time.restype = ctypes.c_long
        done = L.orc_run_to_time(ctypes.byref(self.s), METHODS[method], ctypes.c_double(eps_abs),
                                 ctypes.c_double(eps_rel), ctypes.c_double(h0), ctypes.c_double(t_stop),
                                 ctypes.c_long(int(max_steps)), _p(M), _p(stats))
        return M, dict(t=stats[0], h=stats[1], h_next=stats[2], attempts=int(stats[3]),
                       rejections=int(stats[4]), steps=int(done))

    def run_until(self, M, t_end, method="rkf45", max_steps=1000000, eps_abs=1e-4, eps_rel=1e-4, h0=1e-13):
        """Accepted steps until t_end, the last one clipped (runge_kutta.py:74-78)."""
        M = _c(M).copy()
        stats = np.zeros(5)
        L = lib()
        L.orc_run_until.restype = ctypes.c_long
        done = L.orc_run_until(ctypes.byref(self.s), METHODS[method], ctypes.c_double(eps_abs),
                               ctypes.c_double(eps_rel), ctypes.c_double(h0), ctypes.c_double(t_end),
                               ctypes.c_long(int(max_steps)), _p(M), _p(stats))
        return M, dict(t=stats[0], h=stats[1], h_next=stats[2], attempts=int(stats[3]),
                       rejections=int(stats[4]), steps=int(done))
