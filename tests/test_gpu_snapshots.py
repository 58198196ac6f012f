"""Snapshot writers of the drop-in layers (the reference's snapShootSlice /
snapShootVolum text files, ..._concise/ElasticWaveEquation.cu:929-1037, and a
raw binary volume): the C++ header and the Python mirror write byte-identical
files for the same run, and the files hold the solver's fields."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_and_python_snapshots_agree(tmp_path):
    import sgfd

    lib_dir = os.path.join(ROOT, "staggered-grid_amd", "sgfd")
    exe = tmp_path / "snapshots"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "staggered-grid_amd", "examples", "snapshots.cpp"), "-L" + lib_dir, "-lsgfd",
                    "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    dcpp, dpy = tmp_path / "cpp", tmp_path / "py"
    dcpp.mkdir()
    dpy.mkdir()
    print("[gpu] C++ snapshot program", flush=True)
    r = subprocess.run([str(exe), str(dcpp)], check=True, capture_output=True, text=True, timeout=120)
    assert "snapshots written" in r.stdout

    nx, ny, nz, nt = 80, 70, 60, 41
    n = nx * ny * nz
    med = (np.full(n, 1 / np.float32(2600.0), np.float32), np.full(n, np.float32(2600.0 * 5000.0 * 5000.0), np.float32),
           np.full(n, np.float32(2600.0 * 5000.0 * 5000.0 / 3.0), np.float32))
    dc = [1.21124, -8.97217e-2, 1.38428e-2, -1.76566e-3, 1.18680e-4]
    e = sgfd.ElasticWaveEquation(nt, None, 8, (nx, ny, nz), (0.0005, 10.0, 10.0, 10.0), med, (40, 35, 30),
                                 (50, 40, 35), dc, 20.0)
    try:
        print("[gpu] Python mirror", flush=True)
        e.run(snap_every=20, snap_from=20, snap_dir=str(dpy))
        e.snapShootVolum(nt - 1, str(dpy))
        e.snapShootBinary(nt - 1, "vz", str(dpy))
        F = {f: e.solver.field(f) for f in ("vx", "sxx", "sxy", "vz")}
    finally:
        e.close()

    names = sorted(os.listdir(dcpp))
    assert names == sorted(os.listdir(dpy))
    assert len(names) == 2 * 9 + 2
    for name in names:
        assert filecmp.cmp(dcpp / name, dpy / name, shallow=False), name

    # contents: the last slices and volumes against the fields (text has 6
    # significant digits)
    sx, sy, sz = 40, 35, 30
    for stem, f in (("velocityX", "vx"), ("stressXX", "sxx"), ("stressXY", "sxy")):
        for plane, ref in (("YoZ", F[f][:, :, sx]), ("XoZ", F[f][:, sy, :]), ("XoY", F[f][sz, :, :])):
            got = np.loadtxt(dpy / f"{stem}SliceXX_{plane}_40.txt")
            assert got.shape == (ref.size,)
            assert np.allclose(got, ref.ravel(), rtol=1e-5, atol=1e-30 + 1e-5 * np.abs(ref).max())
    vol = np.loadtxt(dpy / "snapshot_0_40.txt")
    ref = F["sxx"][5:-5, 5:-5, 5:-5].ravel()
    assert np.allclose(vol, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    assert np.array_equal(np.fromfile(dpy / "vz_40.bin", "<f4"), F["vz"].ravel())
