"""CPU-only checks of the C ABI library: it loads, exports every entry point
declared in include/sgfd.h, and its host numerics agree with the oracle.
No device work is started here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "sgfd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_ ]+\*?\s*\**\s*(sg_[a-z0-9_]+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def sgfd_mod():
    import sgfd

    if not os.path.exists(sgfd.LIB_PATH):
        pytest.fail("libsgfd.so is not built: run __graft_entry__.build()")
    sgfd.lib()
    return sgfd


def test_library_exports_every_declared_symbol(sgfd_mod):
    syms = _declared_symbols()
    assert len(syms) >= 30
    L = ctypes.CDLL(sgfd_mod.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert sorted(sgfd_mod.API) == syms


def test_api_version(sgfd_mod):
    assert sgfd_mod.lib().sg_api_version() == 2


def test_ricker_bitexact_with_oracle(sgfd_mod, oracle_mod):
    for f0, dt, mode in ((15.0, 1e-3, "none"), (30.0, 5e-4, "ceil"), (20.0, 5e-4, "ceil")):
        a = sgfd_mod.ricker(f0, dt, 400, mode)
        b = oracle_mod.ricker(f0, dt, 400, oracle_mod.RICKER_CEIL_DELAY if mode == "ceil" else oracle_mod.RICKER_NO_DELAY)
        assert np.array_equal(a, b)


def test_damping_profiles_match_oracle(sgfd_mod, oracle_mod):
    for n, pml, h, kind, v in ((100, 10, 10.0, "quadratic", 3000.0), (60, 7, 5.0, "reference", 2000.0)):
        a = sgfd_mod.damping_profile(n, pml, h, kind, v)
        b = oracle_mod.damping_profile(n, pml, h, oracle_mod.DAMP_QUADRATIC if kind == "quadratic"
                                       else oracle_mod.DAMP_REFERENCE, v, dtype=np.float64)
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=0)


def test_diff_coef(sgfd_mod, oracle_mod):
    np.testing.assert_array_equal(sgfd_mod.diff_coef("reference"), np.array(oracle_mod.REF_COEF, np.float32))
    np.testing.assert_array_equal(sgfd_mod.diff_coef("exact"), np.array(oracle_mod.EXACT_COEF, np.float32))


@pytest.mark.parametrize("kw", [
    dict(nx=10),                 # extent below the stencil minimum
    dict(pml=60),                # PML wider than the grid allows
    dict(dt=0.0),
    dict(precision="fp8"),
])
def test_create_rejects_bad_config_without_gpu(sgfd_mod, kw):
    args = dict(nx=40, ny=40, nz=40, pml=5, dt=1e-3, dx=10.0, dy=10.0, dz=10.0)
    prec = kw.pop("precision", "fp32")
    args.update(kw)
    if prec not in sgfd_mod.PRECISIONS:
        with pytest.raises(KeyError):
            sgfd_mod.Solver(**args, precision=prec)
        return
    with pytest.raises(sgfd_mod.SgError) as e:
        sgfd_mod.Solver(**args)
    assert e.value.code == -1


def test_host_numerics_reject_bad_arguments(sgfd_mod):
    L = sgfd_mod.lib()
    out = np.zeros(8, np.float64)
    assert L.sg_damping_profile(8, 2, 10.0, 1, 3000.0, 1e-3, 3.4, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == -1
    assert b"invalid" in L.sg_last_error()


def _build_example(tmp_path):
    import subprocess

    exe = tmp_path / "main_concise"
    lib_dir = os.path.join(ROOT, "staggered-grid_amd", "sgfd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "staggered-grid_amd", "examples", "main_concise.cpp"),
                    "-L" + lib_dir, "-lsgfd", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    return exe


def test_cpp_dropin_header_compiles_and_links(sgfd_mod, tmp_path):
    """The reference main.cu, ported to include/ElasticWaveEquation.hpp, builds
    with a plain C++ compiler against the C ABI (no HIP headers needed)."""
    exe = _build_example(tmp_path)
    assert exe.exists()


def test_cpu_class_dropin_header_compiles_and_links(sgfd_mod, tmp_path):
    """The reference CPU main.cpp, ported to include/ElasticWaveEquation3D.hpp
    (the CPU driver class: interior GRID, float*** models), builds with a
    plain C++ compiler against the C ABI."""
    import subprocess

    exe = tmp_path / "main_cpu"
    lib_dir = os.path.join(ROOT, "staggered-grid_amd", "sgfd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "staggered-grid_amd", "examples", "main_cpu.cpp"),
                    "-L" + lib_dir, "-lsgfd", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    assert exe.exists()


C_PROGRAM = r"""#include "sgfd.h"
#include <stdio.h>
int main(void)
{
    float c[5];
    int rc = sg_diff_coef(1, c);
    printf("%d %d %.6f\n", sg_api_version(), rc, c[0]);
    return rc;
}
"""


def test_c_header_is_plain_c99(tmp_path):
    """include/sgfd.h is the boundary an FFI binds: it compiles as C99 with
    -pedantic, and a C program links against the library."""
    import subprocess

    src = tmp_path / "abi.c"
    src.write_text(C_PROGRAM)
    lib_dir = os.path.join(ROOT, "staggered-grid_amd", "sgfd")
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    str(src), "-L" + lib_dir, "-lsgfd", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=60).stdout.split()
    assert out[0] == "2" and out[1] == "0" and abs(float(out[2]) - 19845 / 16384) < 1e-6
