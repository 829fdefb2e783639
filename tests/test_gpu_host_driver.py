"""End to end through the host C++ layer: cmg_driver runs a reference-style XML
settings file (Tetgen mesh, Guccione, static pressure ramp) and must reproduce the
per-step results of the same problem driven through the Python host layer."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from cardiomechanics_amd import materials as MT
from cardiomechanics_amd import mesh as M
from cardiomechanics_amd import solver as SV
from cardiomechanics_amd.context import MechanicsContext

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cardiomechanics_amd", "cmg_driver")


def test_cmg_driver_cube_static_matches_python_host(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("cmg_driver not built")
    mesh = M.cube_mesh(2)
    M.write_tetgen(str(tmp_path / "cube"), mesh)
    shutil.copy(os.path.join(ROOT, "cardiomechanics_amd", "configs", "cube_static.xml"), tmp_path / "cube.xml")
    out = subprocess.run([DRIVER, "cube.xml", "-parameter", "Solver.StopTime=0.3", "-parameter",
                          "Solver.Precision=1e-11"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    vols = [float(v) for v in re.findall(r"volume=([-0-9.e+]+)", out.stdout)]
    assert len(vols) == 4          # t = 0, 0.1, 0.2, 0.3
    # the same run through the Python host layer
    with MechanicsContext(mesh, [MT.cube_guccione()], newton_atol=1e-11, newton_rtol=1e-11,
                          newton_stol=1e-11) as ctx:
        recs = SV.run(ctx, "static", SV.Timing(0.0, 0.3, 0.1), SV.linear_ramp(2000.0, 0.0, 1.0))
    ref = [r.volumes[0] for r in recs]
    np.testing.assert_allclose(vols, ref, rtol=1e-9)
    assert os.path.exists(tmp_path / "cube_pressure.dat")
