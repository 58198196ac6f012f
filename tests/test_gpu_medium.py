"""Random media on the GPU (sg_random_medium) against the numpy oracle
(oracle/medium.py) on the same seeds.

Tolerance: the GPU transform is single precision (hipFFT C2C) and the
oracle's double precision from the same complex64 spectrum, so delta is held
to rel-L2 <= 1e-5; V = v0 (1 + delta) then agrees to ~1e-7 relative.  At
512x512x384 the oracle's FFT would take minutes, so that size is checked
through properties that do not depend on the FFT's rounding: zero mean and
standard deviation eps, and seed determinism.
"""
import sys
import numpy as np
import pytest

from conftest import say

import medium as M

pytestmark = pytest.mark.gpu


def _say(msg):
    say(msg)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def sgfd_mod():
    import sgfd

    sgfd.lib()
    return sgfd


CASES = [
    # the script's own spectrum and parameters (a = b = c = 10, p = 10, l = 50)
    dict(nx=64, ny=60, nz=50, kind="reference", a=(10.0, 10.0, 10.0), order=10.0, l=50.0, seed=1),
    # a broader reference-form spectrum, mixed-radix sizes
    dict(nx=48, ny=40, nz=36, kind="reference", a=(3.0, 4.0, 5.0), order=2.0, l=6.0, seed=2),
    dict(nx=48, ny=40, nz=36, kind="gaussian", a=(3.0, 2.0, 4.0), order=0.0, l=1.0, seed=3),
    dict(nx=45, ny=33, nz=28, kind="von_karman", a=(5.0, 5.0, 2.5), order=0.3, l=1.0, seed=4),
    dict(nx=128, ny=96, nz=80, kind="von_karman", a=(8.0, 8.0, 8.0), order=0.5, l=1.0, seed=5),
]


@pytest.mark.parametrize("c", CASES, ids=[f"{c['kind']}-{c['nx']}x{c['ny']}x{c['nz']}" for c in CASES])
def test_random_medium_matches_oracle(sgfd_mod, c):
    _say(f"random medium {c}")
    args = {k: c[k] for k in ("kind", "a", "order", "l", "seed")}
    d = sgfd_mod.random_medium(c["nx"], c["ny"], c["nz"], v0=0.0, eps=0.1, **args)
    want = M.random_medium(c["nx"], c["ny"], c["nz"], v0=0.0, eps=0.1, **args)
    e = _rel(d, want)
    _say(f"delta rel-L2 {e:.2e}")
    assert e <= 1e-5
    v = sgfd_mod.random_medium(c["nx"], c["ny"], c["nz"], v0=2000.0, eps=0.1, **args)
    np.testing.assert_allclose(v, 2000.0 * (1 + want), rtol=2e-6)



SMALL = [
    dict(nx=7, ny=3, nz=1, kind="reference", a=(4.7, 4.5, 5.3), order=8.2, l=7.3, seed=11),
    dict(nx=6, ny=1, nz=2, kind="von_karman", a=(4.9, 5.3, 2.3), order=0.35, l=45.8, seed=12),
    dict(nx=32, ny=2, nz=3, kind="gaussian", a=(11.7, 3.3, 7.9), order=1.75, l=29.0, seed=13),
    dict(nx=2, ny=2, nz=1, kind="reference", a=(10.0, 2.7, 8.1), order=8.8, l=17.0, seed=14),
    # the two fastest axes of one point
    dict(nx=1, ny=1, nz=3, kind="gaussian", a=(1.0, 1.0, 0.8), order=1.0, l=1.0, seed=17),
    dict(nx=2, ny=1, nz=5, kind="gaussian", a=(0.9, 1.0, 1.1), order=1.0, l=1.0, seed=18),
    dict(nx=1, ny=5, nz=4, kind="von_karman", a=(3.0, 2.0, 4.0), order=0.5, l=1.0, seed=15),
    dict(nx=1, ny=1, nz=40, kind="reference", a=(3.0, 3.0, 3.0), order=2.0, l=6.0, seed=16),
]


@pytest.mark.parametrize("c", SMALL, ids=[f"{c['kind']}-{c['nx']}x{c['ny']}x{c['nz']}" for c in SMALL])
def test_random_medium_degenerate_axes(sgfd_mod, c):
    """Axes of one to three points: the spectral method still matches the
    oracle (delta rel-L2 <= 1e-5)."""
    args = {k: c[k] for k in ("kind", "a", "order", "l", "seed")}
    d = sgfd_mod.random_medium(c["nx"], c["ny"], c["nz"], v0=0.0, eps=0.1, **args)
    want = M.random_medium(c["nx"], c["ny"], c["nz"], v0=0.0, eps=0.1, **args)
    assert _rel(d, want) <= 1e-5


@pytest.mark.parametrize("a,seed", [((5.7, 11.5, 11.3), 751968576823), ((2.70, 2.75, 5.14), 777253584999)])
def test_random_medium_refuses_a_flat_field(sgfd_mod, a, seed):
    """Correlation lengths far beyond a 1x1x3 grid (Gaussian spectrum): the
    non-constant modes are 0 or 5e-7 of the constant one, the FP64 field's
    spread is 0 or 9e-7 of its mean.  Single precision holds each value to
    ~6e-8 of the mean, so the library refuses (spread below 1e-4 of the mean)
    instead of scaling rounding noise up to eps; the oracle divides by the
    spread there."""
    raw = M.raw_field(1, 1, 3, "gaussian", a=a, order=7.8, l=47.4, seed=seed)
    assert raw.std() < 1e-6 * abs(raw.mean())
    with pytest.raises(sgfd_mod.SgError, match="spread"):
        sgfd_mod.random_medium(1, 1, 3, "gaussian", a=a, order=7.8, l=47.4, seed=seed, v0=0.0, eps=0.1)


def test_random_medium_feeds_the_solver(sgfd_mod):
    """A GPU-built Vp model, written and read back as V.dat, drives a short
    run: the fields stay finite and the source radiates."""
    import os
    import tempfile

    from sgfd import configs

    n, pml = 64, 8
    _say("random medium -> V.dat -> solver, 64^3")
    vp = sgfd_mod.random_medium(n, n, n, "reference", a=(4.0, 4.0, 4.0), order=2.0, l=8.0, seed=3, v0=3000.0,
                                eps=0.05)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "V.dat")
        sgfd_mod.write_vdat(p, vp)
        vp2 = sgfd_mod.read_vdat(p, n, n, n)
    assert np.array_equal(vp, vp2)
    b, lam, mu = configs.elastic_from_velocity(vp2, vp2 / np.sqrt(3.0), np.full_like(vp2, 2000.0))
    s = sgfd_mod.Solver(n, n, n, pml, 1e-3, 10.0, 10.0, 10.0, time_length=60)
    try:
        s.set_medium(b, lam, mu)
        s.set_damping(*[sgfd_mod.damping_profile(n, pml, 10.0, "quadratic", 3000.0)] * 3)
        s.set_source(n // 2, n // 2, n // 2, sgfd_mod.ricker(15.0, 1e-3, 60, "none"))
        s.step(60)
        s.synchronize()
        vz = s.field("vz")
    finally:
        s.close()
    assert np.all(np.isfinite(vz)) and np.abs(vz).max() > 0


def test_cpp_random_medium_workflow(sgfd_mod, tmp_path):
    """examples/random_medium.cpp: the script's model built through the C ABI,
    written as V.dat, read back and run through the drop-in class; the file
    equals the Python call with the same seed and its statistics are v0 and
    eps * v0."""
    import os
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib_dir = os.path.join(root, "staggered-grid_amd", "sgfd")
    exe = tmp_path / "random_medium"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(root, "include"),
                    os.path.join(root, "staggered-grid_amd", "examples", "random_medium.cpp"), "-L" + lib_dir,
                    "-lsgfd", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    _say("C++ random-medium program")
    r = subprocess.run([str(exe), str(tmp_path)], check=True, capture_output=True, text=True, timeout=120)
    assert "run done" in r.stdout
    v = sgfd_mod.read_vdat(str(tmp_path / "V.dat"), 96, 96, 96)
    assert np.array_equal(v, sgfd_mod.random_medium(96, 96, 96, "reference", seed=7, v0=2000.0, eps=0.1))
    v64 = v.astype(np.float64)
    assert abs(v64.mean() - 2000.0) < 1e-3 and abs(v64.std() - 200.0) < 1e-3
    rows = open(tmp_path / "stationData.txt").read().split()
    assert len(rows) == 200 and np.all(np.isfinite(np.array(rows, np.float64)))
