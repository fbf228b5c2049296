"""The shared-divisor quotient of the device kernels equals IEEE division.

Divisions that share a divisor (HLLE and HLLC states, WENO weights, the face
frame, gamma^ij) or divide by a constant (MP5) are computed by grhd::quot
from the correctly rounded reciprocal: one multiply and two FMAs that give the
correctly rounded quotient.  tests/quotient_check.cpp runs the header's own
function on random and structured FP64 inputs and checks the same three
operations exhaustively at binary32.  Host build with g++ only.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_quotient_equals_ieee_division(tmp_path):
    exe = tmp_path / "quotient_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT / 'grhd_amd' / 'csrc'}", str(ROOT / "tests" / "quotient_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert " 0 mismatches" in r.stdout
