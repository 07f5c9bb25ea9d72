"""The reference-side binding (micromagnum_amd/magneto_amd.py): data layout
between the reference's matrix objects and the library's arrays, on the CPU
(no library call).  The objects are stand-ins with the reference's
Python-visible matrix API (tests/refapi_fakes.py)."""
import numpy as np
import pytest

import refapi_fakes as F
from micromagnum_amd import magneto_amd as B


def _field(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    return rng.standard_normal((3, nz, ny, nx))


@pytest.mark.parametrize("shape", [(5, 3, 2), (1, 4, 3), (7, 1, 1)])
def test_vector_planes_are_soa_x_fastest(shape):
    nx, ny, nz = shape
    soa = _field(shape, 1)  # (3, nz, ny, nx): the library's layout
    vm = F.VectorMatrix(shape, soa.reshape(3, -1))
    assert np.array_equal(B.vector_planes(vm), soa.ravel())
    # cell (x, y, z) component c sits at c*n + x + nx*(y + ny*z)
    v = B.vector_planes(vm).reshape(3, nz, ny, nx)
    x, y, z = nx - 1, ny // 2, nz - 1
    assert tuple(v[:, z, y, x]) == vm.get(x + nx * (y + ny * z))


@pytest.mark.parametrize("shape", [(5, 3, 2), (4, 4, 1)])
def test_to_numpy_view_converts_with_transpose_3210(shape):
    # the reference's to_numpy() is indexed [x, y, z, c]; its SoA form is
    # transpose(3, 2, 1, 0) -- a plain ascontiguousarray gives AoS with the
    # component fastest, which the library would misread
    soa = _field(shape, 2)
    vm = F.VectorMatrix(shape, soa.reshape(3, -1))
    a = vm.to_numpy()
    x, y, z = 1, 2, shape[2] - 1
    assert tuple(a[x, y, z]) == vm.get(x + shape[0] * (y + shape[1] * z))
    assert np.array_equal(B.soa_from_to_numpy(a), soa)
    wrong = np.ascontiguousarray(a).ravel()
    assert not np.array_equal(wrong, soa.ravel())


def test_tensor_matrix_values_are_component_fastest():
    ex, ey, ez = 4, 3, 2
    rng = np.random.default_rng(3)
    N = rng.standard_normal((ez, ey, ex, 6))  # mm_generate_demag_tensor order
    m = F.Matrix((6, ex, ey, ez), N.ravel())
    assert B.dims(m) == (6, ex, ey, ez)
    assert np.array_equal(B.matrix_values(m), N.ravel())


@pytest.mark.parametrize("bulk", [False, True])
def test_store_vector_round_trip(bulk):
    shape = (3, 2, 2)
    soa = _field(shape, 4)
    cls = F.VectorMatrixWithBulkSetter if bulk else F.VectorMatrix
    vm = cls(shape)
    B.store_vector(vm, soa)
    assert np.array_equal(vm.planes(), soa.reshape(3, -1))
    assert vm.set_calls == (0 if bulk else 12)


def test_store_matrix_round_trip():
    m = F.Matrix((4, 3, 1))
    vals = np.arange(12.0) * 1.5
    B.store_matrix(m, vals)
    assert np.array_equal(B.matrix_values(m), vals)


def test_tableau_from_reference_object():
    import oracle as O
    t = O.tableau("dp54")
    tab = B._tableau(F.ButcherTableau(t))
    assert tab.num_steps == 7
    for i in range(7):
        assert tab.a[i] == t["a"][i] and tab.c[i] == t["c"][i] and tab.ec[i] == t["ec"][i]
        for j in range(i):
            assert tab.b[i][j] == t["b"][i][j]


def test_binding_imports_nothing_from_the_package():
    # the file is dropped into the reference tree as src/magnum/magneto_amd.py
    import ast
    import inspect
    tree = ast.parse(inspect.getsource(B))
    mods = set()
    for node in ast.walk(tree):
        if isinstance(node, ast.Import):
            mods.update(a.name.split(".")[0] for a in node.names)
        elif isinstance(node, ast.ImportFrom):
            if node.level:
                # only the reference module it sits beside
                assert node.level == 1 and node.module == "magneto_cpu", node.module
                continue
            mods.add(node.module.split(".")[0])
    assert mods <= {"ctypes", "os", "numpy"}, mods
