import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "staggered-grid_amd"))


_CONFIG = None


def pytest_configure(config):
    global _CONFIG
    _CONFIG = config
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~20 s")


def say(msg):
    """A progress line printed while a test runs (past pytest's output
    capture), so a long GPU test shows that it is working."""
    capman = _CONFIG.pluginmanager.getplugin("capturemanager") if _CONFIG is not None else None
    if capman is None:
        print(f"[gpu] {msg}", flush=True)
        return
    with capman.global_and_fixture_disabled():
        print(f"[gpu] {msg}", flush=True)


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name))
    meta = json.loads(str(z["meta"]))
    return meta, {k: z[k] for k in z.files if k != "meta"}


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as O

    if not os.path.exists(O.LIB_PATH):
        import subprocess

        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return O


def random_case(seed):
    """A seeded random geometry: extents 11 + 2P .. 11 + 2P + 90, P in
    0..12, one of the five box shapes for the interior and the PML boxes, a
    z-chunk of 1..40 planes, the source anywhere in the updated range and
    three stations within 6 cells of it.  In 30 steps the wave travels about
    7 cells; a station much farther out records only the stencil's numerical
    precursor (1e-30 and below, near the bottom of the float range), whose
    relative error carries no meaning."""
    rng = np.random.default_rng(seed)
    pml = int(rng.integers(0, 13))
    dims = [int(rng.integers(11 + 2 * pml, 11 + 2 * pml + 91)) for _ in range(3)]
    shapes = [(4, 1), (8, 1), (8, 2), (16, 1), (16, 2)]
    a, b = (shapes[int(i)] for i in rng.integers(0, 5, 2))
    tile = (a[0], a[1], b[0], b[1], int(rng.integers(1, 41)))
    src = tuple(int(rng.integers(5, n - 5)) for n in dims)
    near = lambda: [min(max(c + int(rng.integers(-6, 7)), 5), n - 6) for c, n in zip(src, dims)]
    return dims, pml, tile, src, np.array([near() for _ in range(3)], np.int32)


def random_workload(seed, nt=40):
    """A random_case geometry as a Workload: a 5 % random P velocity with
    Vs = Vp / 1.8, density 2000, a 20 Hz Ricker, four recorded components
    at the three stations; returns (workload, box shape)."""
    import sgfd
    from sgfd import configs

    (nx, ny, nz), pml, tile, src, st = random_case(seed)
    rng = np.random.default_rng(200 + seed)
    vp = 3000.0 * (1 + 0.05 * rng.standard_normal((nz, ny, nx)))
    b, lam, mu = configs.elastic_from_velocity(vp, vp / 1.8, np.full_like(vp, 2000.0))
    w = sgfd.ricker(20.0, 8e-4, nt, "none")
    wl = configs.Workload(f"r{seed}", nx, ny, nz, pml, 8e-4, 10.0, 10.0, 10.0, nt, b, lam, mu, src, w, st,
                          comps=("vx", "vz", "sxx", "syz"), damp_v=3000.0, vp_max=float(vp.max()), f0=20.0)
    return wl, tile
