"""Matrix / VectorMatrix / Field / VectorField / RectangularMesh with the
reference's Python-visible semantics (src/magneto/matrix/matty.inc.i:46-305,
src/magneto/matrix/matrix/AbstractMatrix.cpp:213-270,
src/magnum/mesh/rectangular_mesh.py:21-31, src/magnum/mesh/field.py).

A matrix is either *uniform* (one value, no storage: the reference's
UNIFORM state), or holds its values on the host (numpy, the reference's
layout: x fastest, SoA components) or on the device (a DeviceArray written
by a libmmhip.so entry point).  Copies between host and device happen on
demand, as the reference's per-device cache does.  Reductions and
normalisation run on the device through the library (no host fallback);
element-wise set-up operations (fill, assign, add, scale, get/set) run where
the data is.
"""
import numpy as np

from . import device as _device

__all__ = ["Shape", "Matrix", "VectorMatrix", "RectangularMesh", "Field", "VectorField"]


def Shape(*dims):
    """Shape(nx, ny, nz) -> tuple (the reference's matty Shape, rank 3 here)."""
    if len(dims) == 1 and isinstance(dims[0], (tuple, list)):
        dims = tuple(dims[0])
    dims = tuple(int(d) for d in dims) + (1,) * (3 - len(dims))
    return dims[:3]


class _Abstract:
    _comps = 1

    def __init__(self, shape):
        self._shape = Shape(shape)
        self._uni = None        # uniform value (float or tuple of 3) or None
        self._host = None       # numpy array in storage order or None
        self._dev = None        # DeviceArray or None
        self._where = "uniform"  # "uniform" | "host" | "device"
        self._uni = self._zero()

    # ---- shape ---------------------------------------------------------
    @property
    def shape(self):
        return self._shape

    @property
    def rank(self):
        return 3

    def size(self):
        nx, ny, nz = self._shape
        return nx * ny * nz

    def dimX(self):
        return self._shape[0]

    def dimY(self):
        return self._shape[1]

    def dimZ(self):
        return self._shape[2]

    def _storage_shape(self):
        nx, ny, nz = self._shape
        return (nz, ny, nx) if self._comps == 1 else (3, nz, ny, nx)

    # ---- state ---------------------------------------------------------
    def isUniform(self):
        return self._where == "uniform"

    def getUniformValue(self):
        if not self.isUniform():
            raise ValueError("%s must be uniform in order to access the uniform value" % type(self).__name__)
        return self._uni

    @property
    def uniform_value(self):
        return self.getUniformValue()

    def _array(self):
        """Host values in storage order (a copy for uniform matrices)."""
        if self._where == "uniform":
            return self._expand(self._uni)
        if self._where == "device":
            self._host = self._dev.numpy().astype(np.float64)
            self._where = "host"
        return self._host

    def _host_mut(self):
        """Host values, owned and writable in place (leaves uniform/device state)."""
        if self._where == "uniform":
            self._host = self._expand(self._uni)
            self._uni = None
        elif self._where == "device":
            self._host = self._dev.numpy().astype(np.float64)
        self._where = "host"
        return self._host

    def _set_host(self, a):
        self._host = np.ascontiguousarray(a, dtype=np.float64).reshape(self._storage_shape())
        self._where = "host"
        self._uni = None

    def device(self):
        """A float64 DeviceArray holding the current values."""
        if self._where == "device":
            return self._dev
        if self._dev is None:
            self._dev = _device.empty(self._storage_shape())
        self._dev.upload(self._array())
        return self._dev

    def device_out(self):
        """A DeviceArray that a library call writes; call mark_device() after."""
        if self._dev is None:
            self._dev = _device.empty(self._storage_shape())
        return self._dev

    def mark_device(self):
        self._where = "device"
        self._uni = None
        self._host = None

    # ---- element-wise set-up -------------------------------------------
    def clear(self):
        self.fill(self._zero())

    def assign(self, other):
        if other.shape != self.shape or other._comps != self._comps:
            raise ValueError("assign: shape mismatch")
        if other.isUniform():
            self.fill(other._uni)
        else:
            self._set_host(np.array(other._array(), copy=True))

    def scale(self, factor):
        if self.isUniform() and np.ndim(factor) == 0:
            self._uni = self._apply_uniform(lambda v: v * factor)
            return
        a = self._array().copy()
        if np.ndim(factor) == 0:
            a *= factor
        else:
            for c in range(3):
                a[c] *= factor[c]
        self._set_host(a)

    def add(self, op, scale=1.0):
        if op.shape != self.shape or op._comps != self._comps:
            raise ValueError("add: shape mismatch")
        if self.isUniform() and op.isUniform():
            self._uni = self._combine_uniform(self._uni, op._uni, scale)
            return
        self._set_host(self._array() + op._array() * scale)

    def to_numpy(self):
        """Values indexed [x, y, z] (scalar) or [x, y, z, c] (vector), as the
        reference's to_numpy (matty.inc.i:342-380)."""
        a = self._array()
        if self._comps == 1:
            return np.ascontiguousarray(a.transpose(2, 1, 0))
        return np.ascontiguousarray(a.transpose(3, 2, 1, 0))

    def from_numpy(self, N):
        N = np.asarray(N, dtype=np.float64)
        if self._comps == 1:
            self._set_host(N.reshape(self._shape).transpose(2, 1, 0))
        else:
            self._set_host(N.reshape(self._shape + (3,)).transpose(3, 2, 1, 0))

    def _index(self, args):
        if len(args) == 1:
            i = int(args[0])
            nx, ny, _ = self._shape
            return i % nx, (i // nx) % ny, i // (nx * ny)
        return tuple(int(v) for v in args)


class Matrix(_Abstract):
    """Scalar field (matty.inc.i:115-188)."""
    _comps = 1

    def _zero(self):
        return 0.0

    def _expand(self, v):
        return np.full(self._storage_shape(), float(v))

    def _apply_uniform(self, f):
        return float(f(self._uni))

    def _combine_uniform(self, a, b, s):
        return float(a + b * s)

    def fill(self, value):
        self._uni = float(value)
        self._where = "uniform"
        self._host = None

    def multiply(self, rhs):
        self._set_host(self._array() * rhs._array())

    def divide(self, rhs):
        self._set_host(self._array() / rhs._array())

    def randomize(self):
        self._set_host(np.random.default_rng().uniform(-1.0, 1.0, self._storage_shape()))

    def maximum(self):
        return float(self._array().max())

    def sum(self):
        return float(self._array().sum())

    def average(self):
        return self.sum() / self.size()

    def get(self, *args):
        if self.isUniform():
            return self._uni
        x, y, z = self._index(args)
        return float(self._array()[z, y, x])

    def set(self, *args):
        *idx, val = args
        if self.isUniform() and self._uni == val:
            return
        x, y, z = self._index(idx)
        self._host_mut()[z, y, x] = val

    def __repr__(self):
        return "Matrix(%r)" % (self.shape,)


class VectorMatrix(_Abstract):
    """Vector field, SoA components (matty.inc.i:190-290)."""
    _comps = 3

    def _zero(self):
        return (0.0, 0.0, 0.0)

    def _expand(self, v):
        nx, ny, nz = self._shape
        a = np.empty((3, nz, ny, nx))
        for c in range(3):
            a[c] = v[c]
        return a

    def _apply_uniform(self, f):
        return tuple(float(f(c)) for c in self._uni)

    def _combine_uniform(self, a, b, s):
        return tuple(float(a[c] + b[c] * s) for c in range(3))

    @property
    def element_size(self):
        return 3

    def fill(self, value):
        self._uni = tuple(float(v) for v in value)
        self._where = "uniform"
        self._host = None

    def randomize(self):
        self._set_host(np.random.default_rng().uniform(-1.0, 1.0, self._storage_shape()))

    def normalize(self, length):
        """M_i <- len_i M_i / |M_i| with the reference CPU norm
        (CPUDevice.cpp:244-271); `length` a number or a Matrix."""
        from . import magneto
        if not isinstance(length, Matrix):
            L = Matrix(self._shape)
            L.fill(float(length))
            length = L
        d = self.device()
        magneto.normalize(d, length.device(), mode=0)
        self.mark_device()

    def absMax(self):
        from . import magneto
        return magneto.abs_max(self.device())

    def dotSum(self, other):
        from . import magneto
        return magneto.dot_sum(self.device(), other.device())

    def average(self):
        from . import magneto
        return tuple(magneto.average(self.device()))

    def sum(self):
        n = self.size()
        return tuple(v * n for v in self.average())

    def maximum(self):
        a = self._array()
        return tuple(float(a[c].max()) for c in range(3))

    def get(self, *args):
        if self.isUniform():
            return self._uni
        x, y, z = self._index(args)
        a = self._array()
        return tuple(float(a[c, z, y, x]) for c in range(3))

    def set(self, *args):
        *idx, val = args
        val = tuple(float(v) for v in val)
        if self.isUniform() and self._uni == val:
            return
        x, y, z = self._index(idx)
        a = self._host_mut()
        for c in range(3):
            a[c, z, y, x] = val[c]

    def __repr__(self):
        return "VectorMatrix(%r)" % (self.shape,)


class RectangularMesh:
    """RectangularMesh(num_nodes, cell_size, periodic_bc="", periodic_repeat=None)
    (src/magnum/mesh/rectangular_mesh.py:21-31): default repeats 1/15/3/2 for
    0/1/2/3 periodic directions."""

    def __init__(self, num_nodes, cell_size, periodic_bc="", periodic_repeat=None):
        self.num_nodes = tuple(int(v) for v in num_nodes)
        self.delta = tuple(float(v) for v in cell_size)
        self.periodic_bc = periodic_bc
        self.pbc = tuple(periodic_bc.find(s) != -1 for s in ("x", "y", "z"))
        if not periodic_repeat:
            periodic_repeat = {0: 1, 1: 15, 2: 3, 3: 2}[sum(1 for b in self.pbc if b)]
        self.periodic_repeat = int(periodic_repeat)

    @property
    def total_nodes(self):
        nx, ny, nz = self.num_nodes
        return nx * ny * nz

    @property
    def cell_volume(self):
        dx, dy, dz = self.delta
        return dx * dy * dz

    @property
    def size(self):
        """(nx dx, ny dy, nz dz), RectangularMesh.getSize (RectangularMesh.h:42)"""
        return tuple(n * d for n, d in zip(self.num_nodes, self.delta))

    def getPosition(self, linidx):
        """centre of the cell with linear index linidx (RectangularMesh.h:45-58)"""
        nx, ny, _ = self.num_nodes
        dx, dy, dz = self.delta
        z, r = divmod(int(linidx), nx * ny)
        y, x = divmod(r, nx)
        return ((x + 0.5) * dx, (y + 0.5) * dy, (z + 0.5) * dz)

    def __repr__(self):
        return "RectangularMesh(%r, %r, periodic_bc=%r, periodic_repeat=%r)" % (
            self.num_nodes, self.delta, self.periodic_bc, self.periodic_repeat)


class Field(Matrix):
    """Scalar field on a mesh (src/magnum/mesh/field.py)."""

    def __init__(self, mesh):
        super().__init__(mesh.num_nodes)
        self.mesh = mesh


class VectorField(VectorMatrix):
    """Vector field on a mesh (src/magnum/mesh/vector_field.py)."""

    def __init__(self, mesh):
        super().__init__(mesh.num_nodes)
        self.mesh = mesh
