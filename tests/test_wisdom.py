"""Launch-shape table ("wisdom"), the counterpart of the reference's FFTW
wisdom file (src/magneto/Magneto.cpp:151-190): file import/export on the host
(no GPU), then on the GPU the tuner and the rule that launch shapes do not
change results."""
import os

import numpy as np
import pytest

from conftest import gpu_available, vortex_field

HEADER = "# libmmhip launch-shape wisdom: mesh precision device lpw_x tk_y tk_z rpw stage_threads"


def _entries(path):
    with open(path) as f:
        return [ln.rstrip("\n") for ln in f if ln.strip() and not ln.startswith("#")]


def test_wisdom_file_round_trip(tmp_path):
    from micromagnum_amd import solver
    solver.wisdom_forget()
    src = tmp_path / "in.wisdom"
    src.write_text("# tuned on one device\n\n"
                   "256x256x16 fp32 gfx950/256 4 8 16 2 256\n"
                   "64x64x4 fp64 gfx950/256 2 4 8 4 512\n")
    assert solver.wisdom_import(str(src)) == 2
    out = tmp_path / "out.wisdom"
    solver.wisdom_export(str(out))
    text = out.read_text().splitlines()
    assert text[0] == HEADER
    assert _entries(str(out)) == ["256x256x16 fp32 gfx950/256 4 8 16 2 256", "64x64x4 fp64 gfx950/256 2 4 8 4 512"]
    # a second import replaces equal keys and keeps the others
    src.write_text("64x64x4 fp64 gfx950/256 1 2 4 2 128\n")
    assert solver.wisdom_import(str(src)) == 1
    solver.wisdom_export(str(out))
    assert _entries(str(out)) == ["256x256x16 fp32 gfx950/256 4 8 16 2 256", "64x64x4 fp64 gfx950/256 1 2 4 2 128"]
    solver.wisdom_forget()
    solver.wisdom_export(str(out))
    assert _entries(str(out)) == []


@pytest.mark.parametrize("bad", [
    "256x256 fp32 gfx950/256 4 8 16 2 256",          # mesh without nz
    "256x256x16 fp16 gfx950/256 4 8 16 2 256",       # unknown precision
    "256x256x16 fp32 gfx950/256 4 8 16 2",           # a value missing
    "256x256x16 fp32 gfx950/256 4 8 16 2 256 7",     # a value too many
    "256x256x16 fp32 gfx950/256 4 -8 16 2 256",      # negative
    "256x256x16 fp32 gfx950/256 4 8 16 2 4096",      # beyond a workgroup
])
def test_invalid_wisdom_file_imports_nothing(tmp_path, bad):
    from micromagnum_amd import solver
    solver.wisdom_forget()
    good = tmp_path / "good.wisdom"
    good.write_text("32x32x1 fp32 gfx950/256 2 4 8 2 128\n")
    solver.wisdom_import(str(good))
    f = tmp_path / "bad.wisdom"
    f.write_text("16x16x16 fp64 gfx950/256 1 1 1 2 64\n" + bad + "\n")
    with pytest.raises(RuntimeError, match="invalid wisdom file"):
        solver.wisdom_import(str(f))
    out = tmp_path / "out.wisdom"
    solver.wisdom_export(str(out))
    assert _entries(str(out)) == ["32x32x1 fp32 gfx950/256 2 4 8 2 128"]
    with pytest.raises(RuntimeError, match="cannot open"):
        solver.wisdom_import(str(tmp_path / "missing.wisdom"))
    solver.wisdom_forget()


def _run(shape, delta, N, M, prec, steps=8):
    from micromagnum_amd.solver import Stepper
    st = Stepper(shape, delta, N=N, precision=prec, method="rkf45", Ms=8.6e5, alpha=0.1, A=1.3e-11,
                 k_uni=1e5, axis1=(0.0, 0.0, 1.0), h_ext=(1e4, 0.0, 0.0), eps_abs=1e-5, eps_rel=1e-5)
    st.set_m(M)
    st.run(steps)
    out = st.get_m(), st.stats(), st.launch_shape()
    st.close()
    return out


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_autotune_records_a_shape_that_gives_the_same_results(tmp_path, prec):
    from micromagnum_amd import magneto as mg
    from micromagnum_amd import solver
    shape, delta = (64, 48, 4), (3e-9, 3e-9, 3e-9)
    N = mg.GenerateDemagTensor(*shape, *delta)
    M = vortex_field(shape, delta, (96e-9, 72e-9), Ms=8.6e5)
    solver.wisdom_forget()
    M0, s0, shape0 = _run(shape, delta, N, M, prec)
    best, ms = solver.autotune(shape, delta, N=N, precision=prec, method="rkf45", Ms=8.6e5, alpha=0.1,
                               A=1.3e-11, k_uni=1e5, axis1=(0.0, 0.0, 1.0), h_ext=(1e4, 0.0, 0.0),
                               eps_abs=1e-5, eps_rel=1e-5, steps=4)
    assert ms > 0.0 and all(v > 0 for v in best.values())
    M1, s1, shape1 = _run(shape, delta, N, M, prec)
    assert shape1 == best
    assert np.array_equal(M1, M0) and s1 == s0
    # an imported entry with other shapes: used, and again the same results
    path = tmp_path / "w"
    solver.wisdom_export(str(path))
    (line,) = [ln for ln in open(path) if not ln.startswith("#")]
    key = " ".join(line.split()[:3])
    other = dict(lpw_x=1, tk_y=4, tk_z=8, rpw=4, stage_threads=512)
    path.write_text("%s %d %d %d %d %d\n" % (key, *(other[f] for f in solver.SHAPE_FIELDS)))
    solver.wisdom_forget()
    assert solver.wisdom_import(str(path)) == 1
    M2, s2, shape2 = _run(shape, delta, N, M, prec)
    assert shape2 == other
    assert np.array_equal(M2, M0) and s2 == s0
    solver.wisdom_forget()


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP GPU")
def test_wisdom_row_count_beyond_one_workgroup_is_cut(tmp_path):
    # an entry whose stage rows would need more than 1024 threads (3 lines
    # per row pair x 64 threads per 512-point line: 24 rows -> 2304 threads)
    # is cut to what one workgroup holds, and the results do not change
    from micromagnum_amd import magneto as mg
    from micromagnum_amd import solver
    shape, delta = (256, 16, 2), (3e-9, 3e-9, 3e-9)
    N = mg.GenerateDemagTensor(*shape, *delta)
    M = vortex_field(shape, delta, (384e-9, 24e-9), Ms=8.6e5)
    solver.wisdom_forget()
    M0, s0, shape0 = _run(shape, delta, N, M, "fp32")
    path = tmp_path / "w"
    # the key (mesh, precision, device) as this GPU writes it: from an export
    solver.autotune(shape, delta, N=N, precision="fp32", method="rkf45", Ms=8.6e5, alpha=0.1, A=1.3e-11,
                    k_uni=1e5, axis1=(0.0, 0.0, 1.0), h_ext=(1e4, 0.0, 0.0), eps_abs=1e-5, eps_rel=1e-5, steps=2)
    solver.wisdom_export(str(path))
    (line,) = [ln for ln in open(path) if not ln.startswith("#")]
    key = " ".join(line.split()[:3])
    path.write_text("%s 0 0 0 24 0\n" % key)
    solver.wisdom_forget()
    assert solver.wisdom_import(str(path)) == 1
    M1, s1, shape1 = _run(shape, delta, N, M, "fp32")
    solver.wisdom_forget()
    assert 2 <= shape1["rpw"] <= 10 and shape1["stage_threads"] <= 1024
    assert np.array_equal(M1, M0) and s1 == s0
