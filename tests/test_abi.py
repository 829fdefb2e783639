"""The C-ABI library loads and exports every symbol include/cmg.h declares (no
device calls here), and the host-side binding matches the header."""
import ctypes
import os
import re

import pytest

from cardiomechanics_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "cmg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cmg_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcmg.so not built (python -m cardiomechanics_amd.build)")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing


def test_struct_layouts_match_header_sizes():
    # cmg_material: 4 int32 + density + 12 params + tension_max + 12 fn doubles
    assert ctypes.sizeof(_lib.CmgMaterial) == 16 + 8 * (1 + 12 + 1 + 12)
    assert ctypes.sizeof(_lib.CmgMesh) == 3 * 8 + 4 * 4 + 8 * 8
    assert ctypes.sizeof(_lib.CmgOptions) == 10 * 4 + 7 * 8 + 4 * 8
    assert ctypes.sizeof(_lib.CmgStepStats) == 4 * 4 + 6 * 8


def test_default_options_are_the_reference_defaults():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcmg.so not built")
    o = _lib.CmgOptions()
    _lib.load().cmg_default_options(ctypes.byref(o))     # pure host function
    assert o.newton_maxit == 50                            # Solver.MaxIterations
    assert o.newton_atol == o.newton_rtol == o.newton_stol == 1e-8   # Solver.Precision
    assert o.newmark_beta == 0.25 and o.newmark_gamma == 0.5          # CBSolverNewmarkBeta.cpp:206-207
    assert o.consistent_mass == 1 and o.fd_epsilon == 1e-12
