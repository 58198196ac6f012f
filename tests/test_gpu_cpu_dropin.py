"""The reference's CPU driver class (3DCPUwithPML_3DArray_GridIncludePMLandOrder/
ElasticWaveEquation3D.h) through its C++ drop-in, include/ElasticWaveEquation3D.hpp,
and the ported main program, staggered-grid_amd/examples/main_cpu.cpp.

* The golden case G0 (interior 12x10x8, pml 3, 160 steps, the reference's
  homogeneous medium and constants, written by the compiled reference):
  curveLine2 samples sxx at (sx, pml+10, sz), which here is the source
  point, fixture station 4.  It must match rel-L2 <= 1e-5 after the
  6-significant-digit printing.  dampingX/Y/Z.txt must hold the fixture's
  damping lines (the reference's damping_a[.][5][5]) expanded over the
  updated box, and zero on the halo.
* A seeded heterogeneous model handed over as [i][j][k] arrays (k fastest),
  300 steps.  The curve files and the snapshots after steps 201, 251 and 300
  must be the same text as a run of the Python mirror on the same model in
  its x-fastest layout, which catches any axis mix-up in the float*** path.
  snapshot_1..5 must exist and be empty, as the reference leaves them.  The
  Python mirror of the class (sgfd.ElasticWaveEquation3D) must write the
  same files byte for byte.
"""
import sys
import os
import subprocess

import numpy as np
import pytest

from conftest import load_golden, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    say(msg)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    import sgfd

    sgfd.lib()
    out = tmp_path_factory.mktemp("cpu_dropin") / "main_cpu"
    lib_dir = os.path.join(ROOT, "staggered-grid_amd", "sgfd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "staggered-grid_amd", "examples", "main_cpu.cpp"),
                    "-L" + lib_dir, "-lsgfd", "-Wl,-rpath," + lib_dir, "-o", str(out)], check=True)
    return out


def _lines(path):
    with open(path) as f:
        return f.read().split("\n")[:-1]


def test_golden_g0_curve_and_damping_files(exe, tmp_path):
    meta, g = load_golden("g0_native_homog.npz")
    L, pml, nt, src = meta["interior"], meta["pml"], meta["nt"], meta["src"]
    assert pml + 10 == src[1] and list(g["stations"][4]) == list(src)
    _say(f"CPU drop-in, golden G0 {L} pml {pml} nt {nt}")
    out = subprocess.run([str(exe), *map(str, L), str(pml), str(nt), *map(str, src)], cwd=tmp_path, check=True,
                         capture_output=True, text=True, timeout=120)
    assert f"curve samples {nt}" in out.stdout
    curve2 = np.loadtxt(tmp_path / "curveLine2.txt")
    assert curve2.shape == (nt,)
    assert _rel(curve2, g["traces"][:, 4, 3]) < 1e-5
    for c in (1, 3):
        assert np.loadtxt(tmp_path / f"curveLine{c}.txt").shape == (nt,)
    assert not list(tmp_path.glob("snapshot_*"))  # none before step 200

    n = [L[a] + 2 * pml + 10 for a in range(3)]
    upd = [np.zeros(m, bool) for m in n]
    for a in range(3):
        upd[a][5:n[a] - 5] = True
    box = upd[0][:, None, None] & upd[1][None, :, None] & upd[2][None, None, :]
    for a, name in enumerate("XYZ"):
        got = np.loadtxt(tmp_path / f"damping{name}.txt").reshape(n)  # i outermost
        line = g["damp_" + "xyz"[a]].astype(np.float64)
        exp = np.where(box, line.reshape([-1 if b == a else 1 for b in range(3)]), 0.0)
        assert np.allclose(got, exp, rtol=1e-5, atol=0), name
        assert np.count_nonzero(got) > 0


def test_heterogeneous_model_matches_python_mirror(exe, tmp_path):
    import sgfd

    L, pml, nt, src = (14, 12, 10), 4, 300, (16, 15, 13)
    nx, ny, nz = (m + 2 * pml + 10 for m in L)
    rng = np.random.default_rng(20181221)
    # Courant number at most ~0.75 (5 m, 0.5 ms, vp <= 3200)
    vp = (2500.0 * (1 + 0.05 * np.clip(rng.standard_normal((nz, ny, nx)), -4, 4))).astype(np.float32)
    rho = (2000.0 + 300.0 * rng.random((nz, ny, nx))).astype(np.float32)
    b = (np.float32(1.0) / rho).astype(np.float32)
    mu = (rho * (vp / np.float32(1.8)) ** 2).astype(np.float32)
    lam = (rho * vp ** 2 - 2 * mu).astype(np.float32)
    for name, arr in (("b", b), ("lambda", lam), ("mu", mu)):
        np.ascontiguousarray(arr.transpose(2, 1, 0)).tofile(tmp_path / f"{name}.raw")  # [i][j][k], k fastest
    _say(f"CPU drop-in, heterogeneous {L} pml {pml} nt {nt}")
    subprocess.run([str(exe), *map(str, L), str(pml), str(nt), *map(str, src), str(tmp_path)], cwd=tmp_path,
                   check=True, capture_output=True, text=True, timeout=120)

    f32 = lambda x: float(np.float32(x))
    s = sgfd.Solver(nx, ny, nz, pml, 0.0005, 5.0, 5.0, 5.0, coef=sgfd.diff_coef("reference"), time_length=nt,
                    source_mode="transient")
    snaps = {}
    try:
        s.set_medium(b, lam, mu)
        s.set_damping(*[sgfd.damping_profile(m, pml, 5.0, "reference", 2000.0, f32(0.001), f32(3.4))
                        for m in (nx, ny, nz)])
        s.set_source(*src, sgfd.ricker(30.0, 0.0005, nt, "ceil"))
        e = pml + 10
        s.set_stations(np.array([[e, src[1], src[2]], [src[0], e, src[2]], [src[0], src[1], e]], np.int32),
                       comps=("sxx",))
        done = 0
        for it in (200, 250, 299):
            s.step(it + 1 - done)
            done = it + 1
            snaps[it + 1] = s.field("sxx")
        tr = s.stations()
    finally:
        s.close()

    tr = tr.reshape(nt, 3)
    for c in range(3):
        assert _lines(tmp_path / f"curveLine{c + 1}.txt") == ["%g" % v for v in tr[:, c]], c
    assert np.abs(tr).max() > 0
    for step, F in snaps.items():
        got = _lines(tmp_path / f"snapshot_0_{step}.txt")
        assert got == ["%g" % v for v in F.transpose(2, 1, 0).ravel()], step  # i outermost, k innermost
        for q in range(1, 6):
            assert os.path.getsize(tmp_path / f"snapshot_{q}_{step}.txt") == 0
    assert sorted(p.name for p in tmp_path.glob("snapshot_0_*")) == ["snapshot_0_201.txt", "snapshot_0_251.txt",
                                                                      "snapshot_0_300.txt"]

    # the Python mirror of the class writes the same files, byte for byte
    py = tmp_path / "py"
    py.mkdir()
    ewq = sgfd.ElasticWaveEquation3D(L, nt, pml, *[a.transpose(2, 1, 0) for a in (b, lam, mu)], src)
    try:
        curves = ewq.propagate(str(py))
    finally:
        ewq.close()
    assert np.array_equal(np.stack(curves, 1), tr)
    names = sorted(p.name for p in py.iterdir())
    assert names == sorted(p.name for p in tmp_path.glob("*.txt")) and len(names) == 3 + 3 + 18
    for name in names:
        assert (py / name).read_bytes() == (tmp_path / name).read_bytes(), name
