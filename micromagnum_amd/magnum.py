"""The reference's user API for the hot path, over the fused device stepper.

Scripts written for the reference's front end (``from magnum import *``)
use these names with the same meaning:

  World, Body, Material (with the Py/Fe/Ni/Co/Au database), Everywhere,
  Cuboid                              src/magnum/micromagnetics/world/*.py
  RectangularMesh                     src/magnum/mesh/rectangular_mesh.py
  StrayField, ExchangeField, AnisotropyField, ExternalField, SpinTorque,
  HomogeneousCurrent                  src/magnum/micromagnetics/*.py
  create_solver                       create_solver.py:27-79
  Solver.solve/step/relax, step handlers
                                      src/magnum/solver/solver.py:76-142,
                                      micro_magnetics_solver.py:37-42
  Condition, EveryNthStep, EveryNthSecond, AfterNthStep, Time
  (TimeGreaterEq), TimeBetween, Relaxed, Always, Never, &, |, ~
                                      src/magnum/solver/condition.py:18-200
  StepHandler, ScreenLog              step_handler.py:19-36, stephandler/

The solver's state holds the model variables the reference's state has
(t, h, step, M, H_tot, dMdt, E_tot and the term energies, deg_per_ns) and
the material parameters of the modules (Ms, alpha, A, k_uniaxial, ...).
Stepping runs on the GPU: M stays in HBM between steps and the solver loop
hands the device as many accepted steps at once as its conditions allow
(a step handler on EveryNthStep(100) gets every 100th state, the device
makes the 99 in between without returning).  Whatever a condition or step
handler reads is evaluated at exactly the states the reference evaluates it
at.

Extensions: create_solver(..., precision="fp32") runs the mixed-precision
bulk path; normalize_mode=1 the FP64 renormalisation.
"""
import math

import numpy as np

from .matty import RectangularMesh  # noqa: F401  (re-exported: the reference's mesh)
from .solver import Stepper

GAMMA = 2.211e5  # src/magnum/micromagnetics/constants.py:33
MU0 = 4.0 * math.pi * 1e-7  # src/magnum/micromagnetics/constants.py:21


# ------------------------------------------------------------ world
class Material(object):
    """A set of material parameters (material.py:20-98)."""

    def __init__(self, params):
        self._params = dict(params)

    def get(self, key):
        return self._params[key]

    def __getattr__(self, key):
        try:
            return self.__dict__["_params"][key]
        except KeyError:
            raise AttributeError(key)

    def __contains__(self, key):
        return key in self._params

    def __repr__(self):
        return "Material(%r)" % (self._params,)

    @staticmethod
    def fromDB(id, **params):
        p = dict(MATERIAL_DB[id])
        p.update(params)
        return Material(p)

    @staticmethod
    def Py(**params):
        return Material.fromDB("Py", **params)

    @staticmethod
    def Fe(**params):
        return Material.fromDB("Fe", **params)

    @staticmethod
    def Ni(**params):
        return Material.fromDB("Ni", **params)

    @staticmethod
    def Co(**params):
        return Material.fromDB("Co", **params)

    @staticmethod
    def Au(**params):
        return Material.fromDB("Au", **params)


# The reference's material database (world/material.py:111-170), values as
# given there.
MATERIAL_DB = {
    "Py": {"id": "Py", "Ms": 8e5, "alpha": 0.01, "A": 13e-12, "sigma": 0.5e7, "amr": 0.1, "P": 0.5, "xi": 0.02},
    "Fe": {"id": "Fe", "Ms": 17e5, "alpha": 0.01, "A": 21e-12, "axis1": (1.0, 0.0, 0.0), "k_uniaxial": 48e3},
    "Ni": {"id": "Ni", "Ms": 4.9e5, "alpha": 0.01, "A": 9e-12, "axis1": (1.0, 0.0, 0.0), "k_uniaxial": -5.7e3},
    "Co": {"id": "Co", "Ms": 14e6, "alpha": 0.5, "A": 30e-12, "axis1": (1.0, 0.0, 0.0), "k_uniaxial": 520e3},
    "Au": {"id": "Au", "Ms": 0, "alpha": 0, "A": 0, "sigma": 4.55e7, "amr": 0},
}


class Shape(object):
    """A region of the mesh: the linear indices of the cells it holds."""

    def getCellIndices(self, mesh):
        nx, ny, nz = mesh.num_nodes
        dx, dy, dz = mesh.delta
        z, y, x = np.meshgrid((np.arange(nz) + 0.5) * dz, (np.arange(ny) + 0.5) * dy, (np.arange(nx) + 0.5) * dx,
                              indexing="ij")
        return np.flatnonzero(self.contains(x, y, z).ravel())

    def contains(self, x, y, z):
        raise NotImplementedError


class Everywhere(Shape):
    """The whole mesh (world/shape.py)."""

    def contains(self, x, y, z):
        return np.ones_like(x, dtype=bool)


class Cuboid(Shape):
    """Cells whose centre lies in the box [p1, p2] (world/cuboid.py)."""

    def __init__(self, p1, p2):
        self.p1 = tuple(min(a, b) for a, b in zip(p1, p2))
        self.p2 = tuple(max(a, b) for a, b in zip(p1, p2))

    def contains(self, x, y, z):
        (a0, a1, a2), (b0, b1, b2) = self.p1, self.p2
        return (x >= a0) & (x <= b0) & (y >= a1) & (y <= b1) & (z >= a2) & (z <= b2)


class Body(object):
    """An id, a material and a shape (world/body.py:20-45)."""

    def __init__(self, id, material, shape=None):
        if not isinstance(id, str) or not isinstance(material, Material):
            raise TypeError("Body(id, material, shape=None)")
        self.id, self.material, self.shape = id, material, shape or Everywhere()

    def __repr__(self):
        return "Body(%r, %r, %r)" % (self.id, self.material, self.shape)


class World(object):
    """A mesh and the bodies in it (world/world.py:25-75): World(mesh,
    Material) or World(mesh, Body, Body, ...)."""

    def __init__(self, mesh, *args):
        if not isinstance(mesh, RectangularMesh):
            raise TypeError("World: The first argument to World must be a RectangularMesh instance")
        if not all(isinstance(x, (Material, Body)) for x in args):
            raise ValueError("World: Invalid/excess parameters in world construction.")
        self.mesh = mesh
        self.bodies = [x for x in args if isinstance(x, Body)]
        mats = [x for x in args if isinstance(x, Material)]
        if len(mats) > 1:
            raise ValueError("World: I can accept at maximum one direct material argument")
        if mats:
            if self.bodies:
                raise ValueError("World: Can't accept direct material argument if there are any Bodies specified")
            self.bodies = [Body("all", mats[0], Everywhere())]

    def findBody(self, body_id):
        for b in self.bodies:
            if b.id == body_id:
                return b
        raise IndexError("No body with id %s found." % body_id)

    def hasBody(self, body_id):
        return any(b.id == body_id for b in self.bodies)


# ------------------------------------------------------------ modules
class _Module(object):
    params = ()
    vector_params = ()

    def __init__(self):
        pass


class LandauLifshitzGilbert(_Module):
    """Ms, alpha (landau_lifshitz_gilbert.py:40-41); added by create_solver."""
    params = ("Ms", "alpha")

    def __init__(self, do_precess=True):
        self.do_precess = do_precess


class StrayField(_Module):
    """Demagnetising field by FFT convolution (stray_field.py:24-55)."""

    def __init__(self, padding=None):
        self.padding = padding


class ExchangeField(_Module):
    params = ("A",)


class AnisotropyField(_Module):
    params = ("k_uniaxial", "k_cubic", "axis1", "axis2")
    vector_params = ("axis1", "axis2")


class ExternalField(_Module):
    """Zeeman field H_ext = offs + amp sin(t freq + phase), or fn(t)
    (external_field.py:21-37, alternating_field.py:24-87)."""
    params = ("H_ext_offs", "H_ext_amp", "H_ext_freq", "H_ext_phase", "H_ext_fn")


class SpinTorque(_Module):
    """Zhang-Li spin torque (spin_torque.py:22-60); the current density is
    the state variable j of HomogeneousCurrent."""
    params = ("xi", "P")

    def __init__(self, do_precess=True):
        self.do_precess = do_precess


class HomogeneousCurrent(_Module):
    """The uniform current density state.j (homogeneous_field.py:89-91)."""

    def __init__(self, default_value=None):
        self.default_value = default_value


# ------------------------------------------------------------ conditions
class Condition(object):
    """condition.py:18-101.  steps_left(state): how many accepted steps
    certainly leave the condition's value unchanged (1 = check every step),
    which lets the solver hand the device several steps at once."""

    def __init__(self, fn=None):
        self.fn = fn

    def check(self, state):
        return self.fn(state)

    def get_time_of_interest(self, state):
        return None

    def steps_left(self, state):
        return 1

    def stop_time(self, state):
        """None, or a time from which on the value may change after a step
        that is not clipped to it: the device returns after the first step
        that ends at or past it (TimeGreaterEq names no time of interest, so
        the reference's solve(Time(t)) overshoots t)."""
        return None

    def __and__(self, other):
        return AndCondition(self, other)

    def __or__(self, other):
        return OrCondition(self, other)

    def __invert__(self):
        return InvertCondition(self)


def _toi(*conds, state):
    t = min([c.get_time_of_interest(state) or 1e100 for c in conds])
    return None if t == 1e100 else t


def _stop(*conds, state):
    ts = [t for t in (c.stop_time(state) for c in conds) if t is not None]
    return min(ts) if ts else None


class AndCondition(Condition):
    def __init__(self, a, b):
        super(AndCondition, self).__init__(lambda s: a.check(s) and b.check(s))
        self._ops = (a, b)

    def get_time_of_interest(self, state):
        return _toi(*self._ops, state=state)

    def steps_left(self, state):
        return min(c.steps_left(state) for c in self._ops)

    def stop_time(self, state):
        return _stop(*self._ops, state=state)


class OrCondition(AndCondition):
    def __init__(self, a, b):
        super(OrCondition, self).__init__(a, b)
        self.fn = lambda s: a.check(s) or b.check(s)


class InvertCondition(Condition):
    def __init__(self, a):
        super(InvertCondition, self).__init__(lambda s: not a.check(s))
        self._a = a

    def get_time_of_interest(self, state):
        return self._a.get_time_of_interest(state)

    def steps_left(self, state):
        return self._a.steps_left(state)

    def stop_time(self, state):
        return self._a.stop_time(state)


class Always(Condition):
    def __init__(self):
        super(Always, self).__init__(lambda s: True)

    def steps_left(self, state):
        return 1 << 40


class Never(Condition):
    def __init__(self):
        super(Never, self).__init__(lambda s: False)

    def steps_left(self, state):
        return 1 << 40


class EveryNthStep(Condition):
    """True at every nth step (condition.py:111-119)."""

    def __init__(self, nth):
        super(EveryNthStep, self).__init__(lambda s: s.step % nth == 0)
        self.nth = int(nth)

    def steps_left(self, state):
        return self.nth - state.step % self.nth


class EveryNthSecond(Condition):
    """True when t is a multiple of nth (condition.py:121-133); asks the
    solver to land on those times."""

    def __init__(self, nth):
        super(EveryNthSecond, self).__init__()
        self.nth = nth

    def check(self, state):
        dt = state.t % self.nth
        return abs(dt) < 1e-16 or abs(dt - self.nth) < 1e-16

    def get_time_of_interest(self, state):
        return (int(state.t / self.nth) + 1) * self.nth

    def steps_left(self, state):
        return 1 << 40  # decided by time, which the solver lands on


class AfterNthStep(Condition):
    def __init__(self, nth):
        super(AfterNthStep, self).__init__(lambda s: s.step > nth)
        self.nth = nth

    def steps_left(self, state):
        return max(1, self.nth + 1 - state.step)


class TimeGreaterEq(Condition):
    """True once t >= t_max (condition.py:145-157).  It names no time of
    interest: steps are not clipped to t_max, the first step that ends at or
    past it makes the condition true."""

    def __init__(self, t_max):
        super(TimeGreaterEq, self).__init__(lambda s: s.t >= t_max)
        self.t_max = t_max

    def steps_left(self, state):
        return 1 if state.t >= self.t_max else 1 << 40

    def stop_time(self, state):
        return self.t_max if state.t < self.t_max else None


Time = TimeGreaterEq


class TimeBetween(Condition):
    """True for t0 <= t < t1 (condition.py:159-167); like TimeGreaterEq it
    names no time of interest, so no step is clipped to t0."""

    def __init__(self, t0, t1):
        super(TimeBetween, self).__init__(lambda s: t0 <= s.t < t1)
        self.t0, self.t1 = t0, t1

    def steps_left(self, state):
        # true on every step in [t0, t1), false for good after t1
        return 1 if self.t0 <= state.t < self.t1 else 1 << 40

    def stop_time(self, state):
        return self.t0 if state.t < self.t0 else None


class Relaxed(Condition):
    """Relaxed(max_degree_per_ns=1, check_every_nth_step=100)
    (condition.py:169-200): at every nth step, deg_per_ns <= the bound and
    below its value at the previous check."""

    def __init__(self, max_degree_per_ns=1, check_every_nth_step=100):
        super(Relaxed, self).__init__()
        self.max_degree_per_ns = max_degree_per_ns
        self.every = int(check_every_nth_step)
        self._last = {}

    def check(self, state):
        if state.step % self.every != 0:
            return False
        d = state.deg_per_ns
        last = self._last.get(id(state))
        self._last[id(state)] = d
        if last is None:
            return False
        return d < last and d <= self.max_degree_per_ns

    def steps_left(self, state):
        return self.every - state.step % self.every


# ------------------------------------------------------------ step handlers
from .stephandler import StepHandler, ScreenLog  # noqa: E402,F401
from .datatable import DataTableLog  # noqa: E402,F401


# ------------------------------------------------------------ state
class _VectorValue(np.ndarray):
    """A host copy of a vector field (3, nz, ny, nx) with the reference's
    VectorMatrix reductions used in scripts (average, absMax, dotSum)."""

    def average(self):
        return tuple(float(v) for v in self.reshape(3, -1).mean(axis=1))

    def absMax(self):
        return float(np.sqrt((np.asarray(self) ** 2).sum(axis=0)).max())

    def dotSum(self, other):
        return float((np.asarray(self) * np.asarray(other)).sum())


def _as_vector(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(_VectorValue)


class _M(object):
    """state.M: reads fetch M from the device; average() uses the device sum."""

    def __init__(self, state):
        self._s = state

    def average(self):
        return tuple(float(v) for v in self._s._stepper_now().average_m())

    def __array__(self, dtype=None, copy=None):
        a = self._s._get_m()
        return a if dtype is None else a.astype(dtype)

    def to_numpy(self):
        return self._s._get_m()


def _float_norm(M):
    """|M| per cell as the reference's CPU normalize3 computes it
    (CPUDevice.cpp:255-259): float components, float sum, float sqrt."""
    e = M.astype(np.float32)
    return np.sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]).astype(np.float64)


class State(object):
    """The solver's state: t, h, step, substep, M and the model variables,
    and the material parameters as attributes."""

    def __init__(self, solver):
        object.__setattr__(self, "_solver", solver)
        self.__dict__.update(t=0.0, h=1e-13, step=0, substep=0)  # micro_magnetics.py:34-39

    # parameters and model variables
    def __getattr__(self, key):
        sv = self.__dict__["_solver"]
        if key in sv._params:
            return sv._params[key]
        if key == "M":
            return _M(self)
        if key in ("H_tot", "H_eff"):
            return _as_vector(sv._device_dmdt()[1])
        if key == "dMdt":
            return _as_vector(sv._device_dmdt()[0])
        if key in ("E_tot", "E_eff", "E_stray", "E_exch", "E_aniso", "E_ext"):
            return sv._device_energy()["E_tot" if key == "E_eff" else key]
        if key == "deg_per_ns":
            return sv._deg_per_ns()
        if key == "j":
            return sv._j
        raise AttributeError(key)

    def __setattr__(self, key, value):
        sv = self._solver
        if key == "M":
            sv._set_m(value)
        elif key == "j":
            sv._set_j(value)
        elif key in sv._params:
            sv._set_param(key, value)
        elif key in ("t", "h"):
            self.__dict__[key] = float(value)
            sv._time_dirty = True
        else:
            self.__dict__[key] = value

    def flush_cache(self):
        self._solver._cache = {}

    def _get_m(self):
        return self._solver._get_m()

    def _stepper_now(self):
        return self._solver._stepper_now()


# ------------------------------------------------------------ solver
class _FieldArg(object):
    """What an assigned function sees as its `field` argument: the field's
    mesh (assign.py:60-65 passes the field being filled)."""

    def __init__(self, mesh):
        self.mesh = mesh


_METHODS = ("rkf45", "rk23", "cc45", "dp54", "rk4", "euler")


class Solver(object):
    """MicroMagneticsSolver over the fused stepper (solver.py:30-142,
    micro_magnetics_solver.py:27-42)."""

    def __init__(self, world, modules, method, eps_abs, eps_rel, step_size, do_precess, precision,
                 normalize_mode):
        self.world, self.mesh = world, world.mesh
        self.modules = modules
        self.method, self.eps_abs, self.eps_rel, self.step_size = method, eps_abs, eps_rel, step_size
        self.do_precess, self.precision, self.normalize_mode = do_precess, precision, normalize_mode
        self._step_handlers = []
        nx, ny, nz = self.mesh.num_nodes
        self._shape = (nz, ny, nx)
        self._params = {}
        self._vector_params = set()
        for m in modules:
            self._vector_params.update(m.vector_params)
        for m in modules:
            for p in m.params:
                self._params[p] = self._default(p)
        self._j = np.zeros((3,) + self._shape)
        for m in modules:
            if isinstance(m, HomogeneousCurrent) and m.default_value is not None:
                self._j = self._vector(m.default_value)
        self._from_world()
        self._M = np.zeros((3,) + self._shape)
        self._stepper = None
        self._dirty = True       # parameters changed: the stepper is rebuilt
        self._m_dirty = True     # M changed on the host
        self._time_dirty = True
        self._cache = {}
        self._N = None
        self._h_next = None
        self.state = State(self)

    # -- parameters
    def _default(self, p):
        if p in self._vector_params:
            return np.zeros((3,) + self._shape)
        if p.startswith("H_ext"):
            return None if p == "H_ext_fn" else (0.0, 0.0, 0.0)
        return np.zeros(self._shape)

    def _vector(self, v):
        if callable(v):
            return self._evaluate(v, 3)
        a = np.asarray(v, dtype=np.float64)
        if a.shape == (3,):
            return np.ascontiguousarray(np.broadcast_to(a.reshape(3, 1, 1, 1), (3,) + self._shape))
        return np.ascontiguousarray(a.reshape((3,) + self._shape))

    def _scalar(self, v):
        if callable(v):
            return self._evaluate(v, 1)
        a = np.asarray(v, dtype=np.float64)
        return np.ascontiguousarray(np.broadcast_to(a, self._shape))

    def _evaluate(self, fn, comps):
        """fn(field, pos) at the cell centres (assign.py:60-65,
        RectangularMesh.getPosition: (i + 0.5) * delta)"""
        nx, ny, nz = self.mesh.num_nodes
        dx, dy, dz = self.mesh.delta
        field = _FieldArg(self.mesh)
        out = np.empty((comps,) + self._shape) if comps > 1 else np.empty(self._shape)
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    v = fn(field, ((i + 0.5) * dx, (j + 0.5) * dy, (k + 0.5) * dz))
                    if comps > 1:
                        out[:, k, j, i] = v
                    else:
                        out[k, j, i] = v
        return out

    def _set_param(self, key, value, mask=None):
        if key.startswith("H_ext"):
            if key == "H_ext_fn":
                self._params[key] = value
            else:
                self._params[key] = tuple(float(x) for x in value)
        else:
            v = self._vector(value) if key in self._vector_params else self._scalar(value)
            if mask is None:
                self._params[key] = v
            else:
                cur = self._params[key].reshape((3, -1) if key in self._vector_params else (-1,)).copy()
                src = v.reshape(cur.shape)
                cur[..., mask] = src[..., mask]
                self._params[key] = cur.reshape(v.shape)
        self._dirty = True
        self._cache = {}

    def _from_world(self):
        """initializeFromWorld (micro_magnetics.py:67-90): each body's
        material parameters on the body's cells."""
        for body in self.world.bodies:
            cells = body.shape.getCellIndices(self.mesh)
            whole = len(cells) == self.mesh.total_nodes
            for p in list(self._params):
                if p in body.material:
                    self._set_param(p, body.material.get(p), None if whole else cells)

    # -- M
    def _set_m(self, value):
        """LandauLifshitzGilbert.update(M) (landau_lifshitz_gilbert.py:83-88):
        assign, then normalise to Ms with the reference's float norm."""
        if isinstance(value, _M):
            value = value.to_numpy()
        if isinstance(value, tuple) and len(value) == 3:
            M = self._vector(value)
        elif callable(value):
            M = self._evaluate(value, 3)
        else:
            M = self._vector(np.asarray(value))
        Ms = self._params["Ms"]
        norm = _float_norm(M)
        with np.errstate(divide="ignore", invalid="ignore"):
            M = np.where(norm == 0.0, M, M * (Ms / np.where(norm == 0.0, 1.0, norm)))
        self._M = np.ascontiguousarray(M)
        self._m_dirty = True
        self._cache = {}

    def _get_m(self):
        if self._stepper is not None and not self._m_dirty:
            self._M = self._stepper.get_m()
        return _as_vector(self._M.copy())

    def _set_j(self, value):
        self._j = self._vector(value)
        self._dirty = True
        self._cache = {}

    # -- the device stepper
    def _spec(self):
        kw = dict(method=self.method, precision=self.precision, normalize_mode=self.normalize_mode,
                  do_precess=self.do_precess, eps_abs=self.eps_abs, eps_rel=self.eps_rel)
        if self.method in ("rk4", "euler"):
            kw["h0"] = self.step_size
        else:
            kw["h0"] = self.state.h
        names = {type(m) for m in self.modules}
        P = self._params
        if ExchangeField in names:
            kw["A"] = P["A"]
        if AnisotropyField in names:
            if np.any(P["k_uniaxial"] != 0):
                kw["k_uni"] = P["k_uniaxial"]
            if np.any(P["k_cubic"] != 0):
                kw["k_cub"] = P["k_cubic"]
            if "k_uni" in kw or "k_cub" in kw:
                kw["axis1"], kw["axis2"] = P["axis1"], P["axis2"]
        if ExternalField in names:
            if P["H_ext_fn"] is not None:
                kw["h_ext_fn"] = P["H_ext_fn"]
            else:
                kw["h_ext"] = P["H_ext_offs"]
                if any(v != 0.0 for v in P["H_ext_amp"]):
                    kw["h_ext_amp"], kw["h_ext_freq"], kw["h_ext_phase"] = (P["H_ext_amp"], P["H_ext_freq"],
                                                                            P["H_ext_phase"])
        if SpinTorque in names:
            kw["P"], kw["xi"], kw["j"] = P["P"], P["xi"], self._j
            st = [m for m in self.modules if isinstance(m, SpinTorque)][0]
            if bool(st.do_precess) != bool(self.do_precess):
                raise NotImplementedError("SpinTorque and the LLG term must agree on do_precess")
        stray = [m for m in self.modules if isinstance(m, StrayField)]
        kw["demag"] = bool(stray)
        if stray:
            if stray[0].padding is not None:
                kw["padding"] = stray[0].padding
            kw["pbc"] = self.mesh.pbc
            kw["periodic_repeat"] = self.mesh.periodic_repeat
            if self._N is not None:
                kw["N"] = self._N
        else:
            kw["pbc"] = self.mesh.pbc
        return kw

    def _stepper_now(self):
        """The device stepper with the current parameters, M, t and h."""
        if self._stepper is None or self._dirty:
            kw = self._spec()
            M = self._get_m() if self._stepper is not None and not self._m_dirty else self._M
            st = Stepper(self.mesh.num_nodes, self.mesh.delta, self._params["Ms"], self._params["alpha"], **kw)
            if kw["demag"] and self._N is None:
                self._N = st.N
            if self._stepper is not None:
                self._stepper.close()
            self._stepper = st
            self._M = np.asarray(M)
            self._m_dirty = True
            self._time_dirty = True
            self._dirty = False
            self._step0 = st.step_count()
        if self._m_dirty:
            self._stepper.set_m(self._M)
            self._m_dirty = False
            self._cache = {}
        if self._time_dirty:
            # the evolver's next trial step: the controller's suggestion
            # (runge_kutta.py:55-58,82), state.h before the first step
            self._stepper.set_time(self.state.t, self._h_next or self.state.h)
            self._time_dirty = False
        return self._stepper

    def _device_dmdt(self):
        if "dmdt" not in self._cache:
            self._cache["dmdt"] = self._stepper_now().dmdt()
        return self._cache["dmdt"]

    def _device_energy(self):
        if "energy" not in self._cache:
            self._cache["energy"] = self._stepper_now().energy()
        return self._cache["energy"]

    def _deg_per_ns(self):
        """landau_lifshitz_gilbert.py:127-131 with h = state.h"""
        dM, _ = self._device_dmdt()
        M = self._get_m()
        dm = float(np.sqrt((np.asarray(dM) ** 2).sum(axis=0)).max())
        mm = float(np.sqrt((np.asarray(M) ** 2).sum(axis=0)).max())
        h = self.state.h
        return 1e-9 * (180.0 / math.pi) * math.atan2(dm * h, mm) / h

    # -- step handlers
    def addStepHandler(self, step_handler, condition):
        if not isinstance(step_handler, StepHandler):
            raise TypeError("'step_handler' argument must be a StepHandler")
        if not isinstance(condition, Condition):
            raise TypeError("'condition' argument must be a Condition")
        if any(s is step_handler for s, _ in self._step_handlers):
            raise ValueError("Cannot add step handler more than once")
        self._step_handlers.append((step_handler, condition))

    def removeStepHandler(self, step_handler):
        for q, (s, _) in enumerate(self._step_handlers):
            if s is step_handler:
                del self._step_handlers[q]
                return
        raise ValueError("step handler not found")

    @property
    def step_handlers(self):
        return self._step_handlers[:]

    def _call_step_handlers(self):
        for h, c in self._step_handlers:
            if c.check(self.state):
                h.handle(self.state)

    # -- the solver loop
    def _advance(self, t_max, steps, t_stop=None):
        """Up to `steps` accepted steps on the device, the last one clipped
        to t_max (runge_kutta.py:74-78), ending early after a step that ends
        at t >= t_stop (not clipped); the state follows."""
        st = self._stepper_now()
        before = st.step_count()
        st.run_until(t_max, max(1, int(steps)), t_stop=t_stop)
        made = st.step_count() - before
        s = st.stats()
        self.state.__dict__.update(t=s["t"], h=s["h"], step=self.state.step + made)
        self._h_next = s["h_next"]
        self._cache = {}
        return made

    def step_with_t_max(self, t_max):
        """solver.py:76-85: one step, clipped to the earliest time of
        interest."""
        t0 = self.state.t
        t1 = min(t for t in [t_max] + [c.get_time_of_interest(self.state) for _, c in self._step_handlers]
                 if t is not None and t > t0)
        self._advance(t1, 1)
        self._call_step_handlers()

    def step(self):
        self.step_with_t_max(1e100)

    def solve(self, stop_condition):
        """solver.py:90-139.  Between two states at which the stop condition
        or a step handler's condition may change, the device makes the steps
        in one call."""
        self.state.flush_cache()
        self._call_step_handlers()
        while not stop_condition.check(self.state):
            t0 = self.state.t
            t1 = min(t for t in [stop_condition.get_time_of_interest(self.state) or 1e100]
                     + [c.get_time_of_interest(self.state) for _, c in self._step_handlers]
                     if t is not None and t > t0)
            n = min([stop_condition.steps_left(self.state)] + [c.steps_left(self.state) for _, c in self._step_handlers])
            t_stop = _stop(stop_condition, *[c for _, c in self._step_handlers], state=self.state)
            self._advance(t1, n, t_stop)
            self._call_step_handlers()
        self.state.flush_cache()

    def relax(self, *args, **kwargs):
        """micro_magnetics_solver.py:37-42: solve(Relaxed(...)).  Without
        step handlers, and when the stepper's own step count lines up with
        state.step, the rule is evaluated on the device (mm_stepper_relax,
        from the k_0 each attempt computes anyway); otherwise by solve()."""
        cond = Relaxed(*args, **kwargs)
        st = self._stepper_now()
        if self._step_handlers or (self.state.step - st.step_count()) % cond.every != 0:
            self.solve(cond)
            return
        before = st.step_count()
        # as solve's loop test, the rule is checked at the current state first
        self.last_relax = st.relax(threshold=cond.max_degree_per_ns, criterion="deg_per_ns",
                                   check_every=cond.every)
        made = st.step_count() - before
        s = st.stats()
        self.state.__dict__.update(t=s["t"], h=s["h"], step=self.state.step + made)
        self._h_next = s["h_next"]
        self._cache = {}


def create_solver(world, module_list=(), **kwargs):
    """create_solver.py:27-79: the LLG module plus the given field modules,
    a Runge-Kutta evolver with MicroMagneticsStepSizeController."""
    do_precess = kwargs.pop("do_precess", True)
    modules = [LandauLifshitzGilbert(do_precess=do_precess)]
    for mod in module_list:
        modules.append(mod() if isinstance(mod, type) else mod)
    evolver = kwargs.pop("evolver", "rkf45")
    if evolver not in _METHODS:
        raise ValueError("Invalid evolver type specified: %s (valid choices: %s)" % (evolver, ", ".join(_METHODS)))
    eps_rel = kwargs.pop("eps_rel", 1e-4)
    eps_abs = kwargs.pop("eps_abs", 1e-3)
    step_size = kwargs.pop("step_size", {"rk4": 3e-13, "euler": 5e-15}.get(evolver))
    precision = kwargs.pop("precision", "fp64")
    normalize_mode = kwargs.pop("normalize_mode", 0)
    log = kwargs.pop("log", False)
    if kwargs:
        raise ValueError("create_solver: Excess parameters given: %s" % ", ".join(kwargs.keys()))
    solver = Solver(world, modules, evolver, eps_abs, eps_rel, step_size, do_precess, precision, normalize_mode)
    if log:
        n = log if type(log) == int else 100
        solver.addStepHandler(ScreenLog(), EveryNthStep(n))
    return solver
