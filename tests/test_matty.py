"""Matrix / VectorMatrix / Field layer (micromagnum_amd.matty) with the
known answers of the reference's src/magnum_tests/vector_matrix_test.py.
Host-side state handling runs on the CPU; reductions and normalisation run
on the device (marked gpu)."""
import numpy as np
import pytest

import oracle as O
from conftest import gpu_available, random_unit_field
from micromagnum_amd.matty import Field, Matrix, RectangularMesh, Shape, VectorField, VectorMatrix



def test_fill_set_get_uniform_state():
    # vector_matrix_test.py:27-38
    m1 = VectorMatrix(Shape(10, 10, 10))
    m1.fill((1.0, 2.0, 3.0))
    assert m1.get(4, 4, 4) == (1.0, 2.0, 3.0)
    assert m1.uniform_value == (1.0, 2.0, 3.0)
    assert m1.isUniform()
    m1.set(4, 4, 4, (3.0, 2.0, 1.0))
    assert m1.get(4, 4, 4) == (3.0, 2.0, 1.0)
    assert m1.get(0, 1, 2) == (1.0, 2.0, 3.0)
    assert not m1.isUniform()
    with pytest.raises(ValueError):
        m1.uniform_value


def test_clear_and_fill():
    # vector_matrix_test.py:40-52
    A = VectorMatrix(Shape(100, 100, 10))
    A.fill((1.0, 2.0, 3.0))
    A.clear()
    assert all(A.get(i) == (0.0, 0.0, 0.0) for i in range(0, A.size(), 997))
    A.fill((1.0, 2.0, 3.0))
    assert all(A.get(i) == (1.0, 2.0, 3.0) for i in range(0, A.size(), 997))


@pytest.mark.parametrize("scale,want", [(1.0, (5, 7, 9)), (2.0, (9, 12, 15))])
def test_add(scale, want):
    # vector_matrix_test.py:54-72
    A = VectorMatrix(Shape(100, 100, 10))
    B = VectorMatrix(Shape(100, 100, 10))
    for idx in range(A.size()):
        A.set(idx, (1 * idx, 2 * idx, 3 * idx))
        B.set(idx, (4 * idx, 5 * idx, 6 * idx))
    A.add(B, scale)
    for idx in range(0, A.size(), 13):
        assert A.get(idx) == tuple(w * idx for w in want)


def test_matrix_ops_and_numpy_layout():
    m = Matrix(Shape(4, 3, 2))
    m.fill(2.0)
    m.scale(1.5)
    assert m.isUniform() and m.uniform_value == 3.0
    m.set(1, 2, 1, 7.0)
    N = m.to_numpy()
    assert N.shape == (4, 3, 2) and N[1, 2, 1] == 7.0 and N[0, 0, 0] == 3.0
    assert m.get(1 + 4 * (2 + 3 * 1)) == 7.0        # linear index x + nx (y + ny z)
    assert m.maximum() == 7.0 and m.sum() == 3.0 * 23 + 7.0
    v = VectorMatrix(Shape(4, 3, 2))
    a = np.random.default_rng(1).standard_normal((4, 3, 2, 3))
    v.from_numpy(a)
    assert np.array_equal(v.to_numpy(), a)
    assert v.get(2, 1, 1) == tuple(a[2, 1, 1])
    w = VectorMatrix(Shape(4, 3, 2))
    w.assign(v)
    w.scale((1.0, 2.0, 3.0))
    assert np.allclose(w.to_numpy(), a * np.array([1.0, 2.0, 3.0]))


def test_mesh_defaults():
    # rectangular_mesh.py:27-31
    assert RectangularMesh((4, 4, 4), (1e-9,) * 3).periodic_repeat == 1
    assert RectangularMesh((4, 4, 4), (1e-9,) * 3, "x").periodic_repeat == 15
    assert RectangularMesh((4, 4, 4), (1e-9,) * 3, "xy").periodic_repeat == 3
    m = RectangularMesh((4, 4, 4), (1e-9,) * 3, "xyz")
    assert m.periodic_repeat == 2 and m.pbc == (True, True, True)
    assert RectangularMesh((4, 4, 4), (1e-9,) * 3, "z", 7).periodic_repeat == 7




def _kat_checks():
    A = VectorMatrix(Shape(10, 10, 10))
    B = VectorMatrix(Shape(10, 10, 10))
    for idx in range(A.size()):
        A.set(idx, (0.1 * idx, 0.2 * idx, -0.3 * idx))
        B.set(idx, (-0.4 * idx, 0.5 * idx, -0.6 * idx))
    assert abs(A.dotSum(B) - 79880040.0) < 0.1
    C = VectorMatrix(Shape(10, 10, 10))
    for idx in range(C.size()):
        C.set(idx, (10, 10, 10))
    C.set(500, (10, 20, 30))
    assert abs(C.absMax() - 37.41657387) < 0.0001
    D = VectorMatrix(Shape(10, 10, 10))
    for idx in range(D.size()):
        D.set(idx, (idx, 2 * idx, 3 * idx))
    assert D.average() == (999.0 / 2, 2 * 999.0 / 2, 3 * 999.0 / 2)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")
def test_reductions_known_answers_on_device():
    _kat_checks()


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")
def test_fields_through_magneto_functions_vs_oracle():
    # exchange(Field Ms, Field A, VectorField M, VectorField H) (exchange.i:35-40)
    # and VectorMatrix.normalize(Matrix) against the oracle
    from micromagnum_amd import magneto as mg
    mesh = RectangularMesh((12, 10, 3), (2e-9, 3e-9, 2e-9), "x")
    Ms, A = Field(mesh), Field(mesh)
    Ms.fill(8e5)
    A.fill(1.3e-11)
    M, H = VectorField(mesh), VectorField(mesh)
    Mn = random_unit_field(mesh.num_nodes, 5) * 8e5
    M.from_numpy(Mn.transpose(3, 2, 1, 0))
    E = mg.exchange(Ms, A, M, H)
    Ho, Eo = O.exchange(mesh.delta, np.full((3, 10, 12), 8e5), np.full((3, 10, 12), 1.3e-11), Mn,
                        pbc=(True, False, False))
    assert np.abs(H.to_numpy().transpose(3, 2, 1, 0) - Ho).max() <= 1e-12 * np.abs(Ho).max()
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    M.scale(1.7)
    M.normalize(Ms)
    got = M.to_numpy().transpose(3, 2, 1, 0)
    want = O.normalize(Mn * 1.7, np.full((3, 10, 12), 8e5), mode=0)
    assert np.array_equal(got, want)
