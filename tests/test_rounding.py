"""Host check of the stochastic rounding used for 16-bit wavefield storage
(no GPU): tests/rounding_check.cpp includes the library's conversion header
(staggered-grid_amd/csrc/sg_precision.hpp) and checks that FP16 and BF16
stochastic rounding return one of the two neighbours, keep exact and
non-finite values, and are unbiased over many draws."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLANG = "/opt/rocm/lib/llvm/bin/clang++"


def test_stochastic_rounding_is_unbiased(tmp_path):
    if not os.path.exists(CLANG):
        pytest.skip("ROCm clang++ not present")
    exe = str(tmp_path / "rounding_check")
    r = subprocess.run([CLANG, "-O2", "-std=c++17", "-x", "c++", "-I", os.path.join(ROOT, "staggered-grid_amd", "csrc"),
                        "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", os.path.join(HERE, "rounding_check.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
