"""The shared-divisor quotient, the division and the square root of the
device kernels equal the IEEE operations.

Divisions that share a divisor (HLLE and HLLC states, WENO weights, the face
frame, gamma^ij) or divide by a constant (MP5) are computed by grhd::quot
from the correctly rounded reciprocal: one multiply and two FMAs that give the
correctly rounded quotient.  tests/quotient_check.cpp runs the header's own
function on random and structured FP64 inputs and checks the same three
operations exhaustively at binary32.  tests/sqrt_div_check.cpp runs the
refinement steps of div64 and sqrt64 (grhd::div_refine, grhd::sqrt_refine),
which the device starts from the v_rcp_f64 and v_rsq_f64 estimates, from
estimates with relative errors up to 2^-14 and checks the IEEE quotient and
square root bit for bit.  Host build with g++ only.
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


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_division_and_square_root_steps_equal_ieee(tmp_path):
    exe = tmp_path / "sqrt_div_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT / 'grhd_amd' / 'csrc'}", str(ROOT / "tests" / "sqrt_div_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "2000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert " 0 mismatches" in r.stdout.splitlines()[-1]
