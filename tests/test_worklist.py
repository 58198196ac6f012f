"""Host check of the tiled sweeps' work-lists (no GPU): the five box lists of
a shot (interior, x/y/z PML face, edges and corners) partition its updated
range exactly, every box damps on exactly the axes of its list, y/z damping
is uniform per box and x tiles are 64-aligned.
The checker (tests/worklist_check.cpp) includes the library's own list
builder, staggered-grid_amd/csrc/sg_worklist.hpp."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("wl") / "worklist_check")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "staggered-grid_amd", "csrc"), "-I/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", os.path.join(HERE, "worklist_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("cannot build the work-list checker: " + r.stderr[-300:])
    return exe


def _check(exe, nx, ny, nz, pml, ty, pty, zc):
    r = subprocess.run([exe] + [str(v) for v in (nx, ny, nz, pml, ty, pty, zc)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("ok"), (nx, ny, nz, pml, ty, pty, zc, r.stdout, r.stderr)
    return tuple(int(v) for v in r.stdout.split()[1:])


def test_benchmark_geometry(checker):
    counts = _check(checker, 512, 512, 384, 20, 16, 8, 12)
    assert len(counts) == 5 and all(c > 0 for c in counts)


@pytest.mark.parametrize("geom", [(28, 26, 24, 3), (150, 77, 61, 3), (100, 100, 100, 10), (64, 64, 80, 6),
                                  (130, 72, 64, 6), (300, 200, 100, 10), (11, 11, 11, 0), (140, 140, 140, 0),
                                  (200, 140, 90, 40), (30, 30, 30, 10), (65, 23, 19, 3), (129, 17, 14, 2),
                                  (37, 90, 21, 4), (75, 12, 80, 1)])
def test_small_and_ragged_geometries(checker, geom):
    for ty, pty, zc in ((4, 4, 1), (8, 4, 5), (16, 8, 13), (16, 16, 64), (4, 16, 200)):
        _check(checker, *geom, ty, pty, zc)


def test_random_geometries(checker):
    rng = np.random.default_rng(7)
    for _ in range(40):
        pml = int(rng.integers(0, 24))
        dims = [int(rng.integers(11 + 2 * pml, 11 + 2 * pml + 150)) for _ in range(3)]
        _check(checker, *dims, pml, int(rng.choice([4, 8, 16])), int(rng.choice([4, 8, 16])),
               int(rng.integers(1, 60)))


def test_gpu_random_parity_geometries(checker):
    """The geometries and box shapes of the GPU random-geometry parity test."""
    from conftest import random_case
    from test_gpu_parity import RANDOM_SEEDS

    for seed in RANDOM_SEEDS:
        (nx, ny, nz), pml, tile, _, _ = random_case(seed)
        _check(checker, nx, ny, nz, pml, tile[0], tile[2], tile[4])


def test_gpu_random_16bit_and_batch_geometries(checker):
    """The geometries and box shapes of the GPU random 16-bit, per-cell
    damping and batch tests (a batch applies one geometry's box shapes to
    three others)."""
    from conftest import random_case

    for seed in (300, 301, 302, 303, 500, 501, 502):
        (nx, ny, nz), pml, tile, _, _ = random_case(seed)
        _check(checker, nx, ny, nz, pml, tile[0], tile[2], tile[4])
    for seed in (400, 401, 402):
        tile = random_case(seed)[2]
        for i in range(3):
            (nx, ny, nz), pml, _, _, _ = random_case(seed * 10 + i)
            _check(checker, nx, ny, nz, pml, tile[0], tile[2], tile[4])
