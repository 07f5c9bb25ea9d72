"""OMF/OVF 1.0 reader and writer (src/magneto/mmm/io/OMFImport.cpp:240-262)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from micromagnum_amd.omf import read_omf, write_omf


def test_reads_reference_fixture_header_and_values():
    f = read_omf(os.path.join(GOLDEN, "M1.omf"))
    assert f.num_nodes == (200, 50, 1)
    assert f.delta == (2.5e-09, 2.5e-09, 3e-09)
    mag = np.sqrt((f.data ** 2).sum(axis=0))
    assert float(f.header["valuerangemaxmag"]) == 800000.0000000004
    assert np.abs(mag - 8e5).max() < 1e-6 * 8e5
    assert f.data[0, 0, 0, 0] == 611043.47683628812


@pytest.mark.parametrize("fmt,tol", [("text", 0.0), ("binary 8", 0.0), ("binary 4", 1e-7)])
def test_round_trip(tmp_path, fmt, tol):
    rng = np.random.default_rng(3)
    data = rng.standard_normal((3, 2, 3, 5)) * 8e5
    p = os.path.join(tmp_path, "f.omf")
    write_omf(p, data, (1e-9, 2e-9, 3e-9), fmt=fmt)
    g = read_omf(p)
    assert g.num_nodes == (5, 3, 2)
    assert np.allclose(g.data, data, rtol=tol, atol=0)
