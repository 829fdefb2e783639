"""Builds and runs the host C++ layer's CPU self-test (settings, time functions,
step control, Tetgen/LAT, material keys); links libcmg.so but makes no device call."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cardiomechanics_amd")


def test_host_layer_selftest(tmp_path):
    if not os.path.exists(os.path.join(PKG, "libcmg.so")):
        pytest.skip("libcmg.so not built")
    exe = tmp_path / "host_selftest"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                           os.path.join(PKG, "host"), os.path.join(ROOT, "tests", "host", "host_selftest.cpp"),
                           os.path.join(PKG, "host", "cm_host.cpp"), os.path.join(PKG, "host", "cm_circulation.cpp"),
                           os.path.join(PKG, "host", "cm_export.cpp"),
                           "-L", PKG, "-lcmg",
                           f"-Wl,-rpath,{PKG}", "-o", str(exe)])
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
