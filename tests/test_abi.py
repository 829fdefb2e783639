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
    assert ctypes.sizeof(_lib.CmgOptions) == 12 * 4 + 7 * 8 + 4 * 8
    assert ctypes.sizeof(_lib.CmgStepStats) == 4 * 4 + 6 * 8 + 2 * 4 + 8
    assert ctypes.sizeof(_lib.CmgBatchInfo) == 8 * 4 + 12 * 8


def test_default_options_are_the_reference_defaults():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcmg.so not built")
    o = _lib.CmgOptions()
    _lib.load().cmg_default_options(ctypes.byref(o))     # pure host function
    assert o.newton_maxit == 50                            # Solver.MaxIterations
    assert o.newton_atol == o.newton_rtol == o.newton_stol == 1e-8   # Solver.Precision
    assert o.newmark_beta == 0.25 and o.newmark_gamma == 0.5          # CBSolverNewmarkBeta.cpp:206-207
    assert o.consistent_mass == 1 and o.fd_epsilon == 1e-12


def test_struct_sizes_agree_with_the_c_compiler(tmp_path):
    """sizeof / offsetof of the header's structs as gcc sees them == the ctypes binding"""
    import shutil
    import subprocess

    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    structs = {"cmg_material": _lib.CmgMaterial, "cmg_mesh": _lib.CmgMesh, "cmg_options": _lib.CmgOptions,
               "cmg_solve_stats": _lib.CmgSolveStats, "cmg_step_stats": _lib.CmgStepStats,
               "cmg_batch_info": _lib.CmgBatchInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cmg.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'    printf("{name} %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'    printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for name, cls in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, f"{name}.{f}"
