"""Fused micromagnetic stepper (device-resident LLG + Runge-Kutta).

Counterpart of the reference's create_solver/Solver.solve loop
(src/magnum/micromagnetics/create_solver.py:27-79, src/magnum/solver/solver.py:76-142)
for the modules StrayField, ExchangeField, AnisotropyField, ExternalField and
SpinTorque, in that order, driven by rkf45/cc45 (MicroMagneticsStepSizeController)
or classical rk4 / euler with a fixed step.  The whole attempt loop runs on the device;
see include/mm_hotpath.h (mm_stepper_*).
"""
import ctypes
import os

import numpy as np

from . import _lib
from .magneto import GenerateDemagTensor, PADDING_ROUND_4


def _field(a, shape, comps=1):
    nx, ny, nz = shape
    if a is None:
        return None
    full = ((comps,) if comps > 1 else ()) + (nz, ny, nx)
    a = np.asarray(a, dtype=np.float64)
    if comps > 1 and a.ndim == 1:
        a = a.reshape(comps, 1, 1, 1)
    return np.ascontiguousarray(np.broadcast_to(a, full))


class SystemSpec:
    """One system's mm_system_desc_t plus the host arrays it points to."""

    def __init__(self, shape, delta, Ms, alpha, A=None, demag=True, N=None, k_uni=None, k_cub=None,
                 axis1=None, axis2=None, h_ext=None, P=None, xi=None, j=None, do_precess=True,
                 method="rkf45", eps_abs=1e-4, eps_rel=1e-4, h0=1e-13, precision="fp64",
                 normalize_mode=0, padding=PADDING_ROUND_4, pbc=(False, False, False), periodic_repeat=None,
                 cache_dir=None, h_ext_amp=None, h_ext_freq=None, h_ext_phase=None, tensor="auto"):
        """h_ext: the Zeeman field (the offset of the reference's
        ExternalField); h_ext_amp, h_ext_freq, h_ext_phase (3-tuples) add
        amp sin(t freq + phase) per component (alternating_field.py:57-84).
        tensor: where GenerateDemagTensor runs when N is not given -- "host",
        "gpu" (open boundaries), or "auto": the GPU for open meshes of at
        least 2^18 cells, the host otherwise."""
        nx, ny, nz = shape
        self.shape = tuple(shape)
        self.delta = tuple(float(v) for v in delta)
        self.n = nx * ny * nz
        self.keep = keep = []

        def f(a, comps=1):
            arr = _field(a, shape, comps)
            if arr is None:
                return None
            keep.append(arr)
            return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        d = _lib.SystemDesc()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.dx, d.dy, d.dz = self.delta
        d.precision = _lib.MM_FP64 if precision == "fp64" else _lib.MM_FP32
        d.method = _lib.METHODS[method]
        d.eps_abs, d.eps_rel, d.h0 = float(eps_abs), float(eps_rel), float(h0)
        d.normalize_mode = int(normalize_mode)
        d.do_precess = int(bool(do_precess))
        d.Ms_host = f(Ms)
        d.alpha_host = f(alpha)
        d.A_host = f(A)
        d.k_uni_host = f(k_uni)
        d.k_cub_host = f(k_cub)
        d.axis1_host = f(axis1, 3)
        d.axis2_host = f(axis2, 3)
        d.P_host = f(P)
        d.xi_host = f(xi)
        d.j_host = f(j, 3)
        if h_ext is not None or h_ext_amp is not None:
            d.use_zeeman = 1
            for c in range(3):
                d.h_ext[c] = float(h_ext[c]) if h_ext is not None else 0.0
        if h_ext_amp is not None:
            d.use_ac = 1
            for c in range(3):
                d.h_amp[c] = float(h_ext_amp[c])
                d.h_freq[c] = float(h_ext_freq[c]) if h_ext_freq is not None else 0.0
                d.h_phase[c] = float(h_ext_phase[c]) if h_ext_phase is not None else 0.0
        d.px, d.py, d.pz = (int(bool(b)) for b in pbc)
        if periodic_repeat is None:
            # RectangularMesh defaults (src/magnum/mesh/rectangular_mesh.py:27-31)
            periodic_repeat = {0: 1, 1: 15, 2: 3, 3: 2}[sum(1 for b in pbc if b)]
        self.pbc = tuple(bool(b) for b in pbc)
        if demag:
            if N is None:
                on_gpu = tensor == "gpu" or (tensor == "auto" and not any(self.pbc) and self.n >= 1 << 18
                                             and cache_dir is None)
                if on_gpu:
                    N = GenerateDemagTensor(nx, ny, nz, *self.delta, padding=padding, device="gpu")
                else:
                    N = GenerateDemagTensor(nx, ny, nz, *self.delta, *self.pbc, repeat=periodic_repeat,
                                            padding=padding, cache_dir=cache_dir)
            N = np.ascontiguousarray(N, dtype=np.float64)
            keep.append(N)
            d.N_host = N.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            d.ez, d.ey, d.ex = N.shape[:3]
        self.N = N if demag else None
        self.desc = d
        self.precision = precision
        self.method = method


SHAPE_FIELDS = ("lpw_x", "tk_y", "tk_z", "rpw", "stage_threads")


def autotune(shape, delta, Ms, alpha, steps=20, **kw):
    """Measure launch shapes for this system (mm_autotune) and record the
    fastest in the process's launch-shape table, which steppers created
    afterwards for the same mesh, precision and device use.  Returns
    (shape dict, ms per step)."""
    kw = dict(kw)
    _zeeman_fn_args(kw)
    spec = SystemSpec(shape, delta, Ms, alpha, **kw)
    best = _lib.LaunchShape()
    ms = ctypes.c_double()
    _lib.check(_lib.device_lib().mm_autotune(ctypes.byref(spec.desc), int(steps), ctypes.byref(best),
                                             ctypes.byref(ms)))
    return {f: getattr(best, f) for f in SHAPE_FIELDS}, ms.value


def wisdom_import(path):
    """Add the entries of a launch-shape file (the reference's FFTW wisdom
    import, Magneto.cpp:151-174); returns their number."""
    n = ctypes.c_int()
    _lib.check(_lib.load().mm_wisdom_import(os.fsencode(path), ctypes.byref(n)))
    return n.value


def wisdom_export(path):
    """Write the launch-shape table (Magneto.cpp:176-190)."""
    _lib.check(_lib.load().mm_wisdom_export(os.fsencode(path)))


def wisdom_forget():
    _lib.check(_lib.load().mm_wisdom_forget())


def _zeeman_fn_args(kw):
    """Take h_ext_fn out of the keyword arguments.  As the reference's
    AlternatingField (alternating_field.py:67-69), a function excludes the
    offset/amplitude/frequency/phase parameters unless they are zero vectors."""
    fn = kw.pop("h_ext_fn", None)
    if fn is None:
        return None
    names = ("h_ext", "h_ext_amp", "h_ext_freq", "h_ext_phase")
    for name in names:
        v = kw.get(name)
        if v is not None and any(float(x) != 0.0 for x in v):
            raise ValueError("If h_ext_fn is defined, the parameters %s must be zero vectors, i.e. (0.0, 0.0, 0.0)"
                             % ", ".join(names))
        kw.pop(name, None)
    return fn


class Stepper:
    def __init__(self, shape, delta, Ms, alpha, **kw):
        """Keyword arguments as SystemSpec, plus h_ext_fn: a function of the
        time t returning the uniform Zeeman field (a 3-vector, A/m), the
        reference's ExternalField with H_ext_fn; it is called on the host at
        every stage time."""
        L = _lib.device_lib()
        kw = dict(kw)
        fn = _zeeman_fn_args(kw)
        spec = SystemSpec(shape, delta, Ms, alpha, **kw)
        self.shape, self.delta, self.n, self.pbc = spec.shape, spec.delta, spec.n, spec.pbc
        self.N, self.precision, self.method = spec.N, spec.precision, spec.method
        h = ctypes.c_void_p()
        _lib.check(L.mm_stepper_create(ctypes.byref(spec.desc), ctypes.byref(h)))
        self._h = h
        self._fn_cb = None
        self._fn_error = None
        if fn is not None:
            self.set_field_fn(fn)

    def launch_shape(self):
        """The launch shape this stepper runs with (see autotune)."""
        r = _lib.LaunchShape()
        _lib.check(_lib.load().mm_stepper_launch_shape(self._h, ctypes.byref(r)))
        return {f: getattr(r, f) for f in SHAPE_FIELDS}

    def set_field_fn(self, fn):
        """Zeeman field from fn(t) -> 3-vector (None removes it)."""
        if fn is None:
            _lib.check(_lib.load().mm_stepper_set_field_fn(self._h, _lib.FIELD_FN(), None))
            self._fn_cb = None
            return

        def cb(t, h, user):
            try:
                v = fn(t)
                v = [float(x) for x in v]
                if len(v) != 3:
                    raise TypeError("h_ext_fn must return a 3-vector (a uniform field)")
                h[0], h[1], h[2] = v
            except BaseException as e:  # reported after the call that evaluated fn
                if self._fn_error is None:
                    self._fn_error = e
                h[0] = h[1] = h[2] = float("nan")

        cb_c = _lib.FIELD_FN(cb)
        _lib.check(_lib.load().mm_stepper_set_field_fn(self._h, cb_c, None))
        self._fn_cb = cb_c

    def _fn_check(self):
        if self._fn_error is not None:
            e, self._fn_error = self._fn_error, None
            raise e

    def _dp(self, a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def set_m(self, M):
        M = _field(M, self.shape, 3).copy()
        _lib.check(_lib.load().mm_stepper_set_m(self._h, self._dp(M)))

    def get_m(self):
        nx, ny, nz = self.shape
        M = np.empty((3, nz, ny, nx))
        _lib.check(_lib.load().mm_stepper_get_m(self._h, self._dp(M)))
        return M

    def dmdt(self):
        nx, ny, nz = self.shape
        dM = np.empty((3, nz, ny, nx))
        H = np.empty((3, nz, ny, nx))
        _lib.check(_lib.load().mm_stepper_dmdt(self._h, self._dp(dM), self._dp(H)))
        self._fn_check()
        return dM, H

    def run(self, steps):
        _lib.check(_lib.load().mm_stepper_run(self._h, int(steps)))
        self._fn_check()

    def run_until(self, t_end, max_steps=10000000, t_stop=None):
        """Step until t == t_end (the last step clipped as in the reference's
        RungeKutta.evolve) or max_steps accepted steps; with t_stop, also
        stop after the first step that ends at t >= t_stop (that step is not
        clipped: the solver loop's stop condition Time(t_stop))."""
        L = _lib.load()
        if t_stop is None:
            _lib.check(L.mm_stepper_run_until(self._h, float(t_end), int(max_steps)))
        else:
            _lib.check(L.mm_stepper_run_until_stop(self._h, float(t_end), float(t_stop), int(max_steps)))
        self._fn_check()

    def sync(self):
        _lib.check(_lib.load().mm_stepper_sync(self._h))

    def stats(self):
        s = np.zeros(5)
        _lib.check(_lib.load().mm_stepper_stats(self._h, self._dp(s)))
        return dict(t=s[0], h=s[1], h_next=s[2], attempts=int(s[3]), rejections=int(s[4]))

    def average_m(self):
        a = np.zeros(3)
        _lib.check(_lib.load().mm_stepper_average_m(self._h, self._dp(a)))
        return a

    def step_count(self):
        """Accepted steps since the stepper was created."""
        n = ctypes.c_longlong()
        _lib.check(_lib.load().mm_stepper_step_count(self._h, ctypes.byref(n)))
        return n.value

    def set_time(self, t, h_next=0.0):
        """Set t and (h_next > 0) the step size of the next attempt."""
        _lib.check(_lib.load().mm_stepper_set_time(self._h, float(t), float(h_next)))

    ENERGY_TERMS = ("E_stray", "E_exch", "E_aniso", "E_ext", "E_tot")

    def energy(self):
        """Energies of the terms at the current M, in joule (the reference's
        E_stray, E_exch, E_aniso, E_ext and their sum E_tot)."""
        e = np.zeros(5)
        _lib.check(_lib.load().mm_stepper_energy(self._h, self._dp(e)))
        self._fn_check()
        return dict(zip(self.ENERGY_TERMS, e.tolist()))

    def relax(self, threshold=1e7, criterion="dmdt", check_every=1, max_steps=1000000):
        """Step until relaxed (see mm_stepper_relax).  criterion "dmdt":
        scaled_abs_max(dM/dt, Ms) < threshold [1/s]; "deg_per_ns": the
        reference's Relaxed(threshold, check_every) condition."""
        crit = {"dmdt": _lib.MM_RELAX_DMDT, "deg_per_ns": _lib.MM_RELAX_DEG_PER_NS}[criterion]
        info = np.zeros(3)
        _lib.check(_lib.load().mm_stepper_relax(self._h, crit, float(threshold), int(check_every),
                                                int(max_steps), self._dp(info)))
        self._fn_check()
        return dict(steps=int(info[0]), relaxed=bool(info[1]), value=float(info[2]))

    KERNEL_CLASSES = ("x_fwd", "y_fwd", "z_mult", "y_inv", "y_mult_2d", "x_inv_stage", "controller")

    def kernel_profile(self, attempts=5):
        """Per kernel class: average ms per launch, launches, algorithmic bytes
        per launch, from HIP events recorded between the kernels."""
        out = np.zeros(3 * len(self.KERNEL_CLASSES))
        _lib.check(_lib.load().mm_stepper_kernel_profile(self._h, int(attempts), self._dp(out)))
        res = {}
        for k, name in enumerate(self.KERNEL_CLASSES):
            if out[3 * k + 1] > 0:
                res[name] = dict(ms=out[3 * k], launches=int(out[3 * k + 1]), bytes=out[3 * k + 2])
        return res

    @property
    def stream(self):
        return _lib.load().mm_stepper_stream(self._h)

    def close(self):
        if getattr(self, "_h", None) and _lib._LIB is not None:
            _lib._LIB.mm_stepper_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Independent systems stepped together (mm_batch_*); `specs` is a list of
    keyword dicts for SystemSpec (shape, delta, Ms, alpha, ...)."""

    def __init__(self, specs, nstreams=4):
        L = _lib.device_lib()
        self.specs = [SystemSpec(**kw) for kw in specs]
        arr = (_lib.SystemDesc * len(self.specs))(*[sp.desc for sp in self.specs])
        h = ctypes.c_void_p()
        _lib.check(L.mm_batch_create(arr, len(self.specs), int(nstreams), ctypes.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.specs)

    @staticmethod
    def _dp(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def set_m(self, i, M):
        M = _field(M, self.specs[i].shape, 3).copy()
        _lib.check(_lib.load().mm_batch_set_m(self._h, int(i), self._dp(M)))

    def get_m(self, i):
        nx, ny, nz = self.specs[i].shape
        M = np.empty((3, nz, ny, nx))
        _lib.check(_lib.load().mm_batch_get_m(self._h, int(i), self._dp(M)))
        return M

    def run(self, steps):
        _lib.check(_lib.load().mm_batch_run(self._h, int(steps)))

    def stats(self, i):
        s = np.zeros(5)
        _lib.check(_lib.load().mm_batch_stats(self._h, int(i), self._dp(s)))
        return dict(t=s[0], h=s[1], h_next=s[2], attempts=int(s[3]), rejections=int(s[4]))

    def average_m(self, i):
        a = np.zeros(3)
        _lib.check(_lib.load().mm_batch_average_m(self._h, int(i), self._dp(a)))
        return a

    def close(self):
        if getattr(self, "_h", None) and _lib._LIB is not None:
            _lib._LIB.mm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
