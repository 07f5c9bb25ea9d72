"""The C-ABI library loads without a GPU and exports every symbol that
include/*.h declares; the Python binding declares the same set.  No compute
calls are made here."""
import ctypes
import glob
import os
import re

from conftest import ROOT
from micromagnum_amd import _lib


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_hot_path():
    names = declared_symbols()
    for required in ["mm_exchange", "mm_uniaxial_anisotropy", "mm_cubic_anisotropy", "mm_llge",
                     "mm_fdm_zhangli", "mm_generate_demag_tensor", "mm_conv_create", "mm_conv_execute",
                     "mm_scaled_abs_max", "mm_rk_prepare_step", "mm_rk_combine_result", "mm_stepper_run"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_symbols() == set(_lib.SIGNATURES)


def test_errors_are_reported_without_gpu():
    L = _lib.load()
    out = ctypes.c_int(0)
    rc = L.mm_round_tensor_dimension(0, 0, 2, ctypes.byref(out))
    assert rc == 1 and b"bad argument" in L.mm_last_error()
    rc = L.mm_round_tensor_dimension(10, 0, 99, ctypes.byref(out))
    assert rc != 0 and b"padding" in L.mm_last_error().lower()
