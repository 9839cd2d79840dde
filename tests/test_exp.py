"""Host check of the FP64 exponential of the covariance code
(nonstationary_mra_amd/csrc/mra_exp.hpp, Tang's table-driven method with 64
entries): compiled with g++ and compared with the C library's exp over 4
million points of [-708, 0] and at special arguments."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exp_neg_within_two_ulp(tmp_path):
    exe = str(tmp_path / "exp_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "nonstationary_mra_amd", "csrc"),
                           os.path.join(ROOT, "tests", "exp_check.cpp"), "-o", exe])
    worst, at, underflow = subprocess.check_output([exe], text=True).split()
    print("worst ulp", worst, "at", at)
    assert float(worst) <= 2.0
    assert underflow == "1"  # below the normal range: exactly 0
