"""Structural zeros of the metric and of the HLLC face frame.

The face and cell kernels drop the terms of their metric sums that a
structural zero multiplies (grhd_amd/csrc/grhd_physics.hpp: GRHD_CZERO,
cof_zero, FaceGeo::gam): gamma_{r theta} and gamma_{theta phi} of every metric
kind, and the coframe entries of face_frame (fluxes.f90:402-473) that the
Gram-Schmidt construction leaves at zero when those components vanish.
tests/frame_zero_check.cpp evaluates metric_face and face_frame at random
positions on all eight metric kinds, in FP64 and FP32, and checks that every
entry marked zero is exactly zero and that the entries kept are not (so a
pattern marking too much, or another direction's pattern, fails).  Host build
with g++ only.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_metric_and_frame_structural_zeros(tmp_path):
    exe = tmp_path / "frame_zero_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT / 'grhd_amd' / 'csrc'}", str(ROOT / "tests" / "frame_zero_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "20000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.splitlines()[-1] == "all structural zeros hold"
