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


def _ref_text(values):
    # ostream << float with the default precision: printf's %g
    return ["%g" % float(np.float32(v)) for v in values]


def test_snapshot_text_against_oracle_fields(tmp_path, oracle_mod):
    """The slice and volume files written from a GPU run against text built
    independently from the CPU oracle's fields of the same run, following
    the reference writers' loops (..._concise/ElasticWaveEquation.cu:
    929-1037): YoZ slices iz-major then iy, XoZ iz then ix, XoY iy then ix,
    the sxx volume over [5, n-5) iz, iy, ix; one value per line, 6 significant
    digits.  Same line count and order; every value within 1e-5 of the field
    maximum of the oracle's; "0" on every line of a halo cell; every line
    printf %g text; above 1e-3 of the field maximum, within one unit of
    the 6th printed digit plus 1e-7 of the maximum of the oracle's text."""
    import sgfd

    O = oracle_mod
    nx, ny, nz, nt, pml = 80, 70, 60, 41, 8
    sx, sy, sz = 40, 35, 30
    n = nx * ny * nz
    b = np.full(n, 1 / np.float32(2600.0), np.float32)
    lam = np.full(n, np.float32(2600.0 * 5000.0 * 5000.0), np.float32)
    mu = np.full(n, np.float32(2600.0 * 5000.0 * 5000.0 / 3.0), np.float32)
    dt, h, f0 = 0.0005, 10.0, 20.0
    p = O.OracleParams(nx, ny, nz, pml, dt, h, h, h, nt, (sx, sy, sz), coef=O.REF_COEF, damp_mode=O.DAMP_REFERENCE,
                       damp_v=2000.0, inject_mode=O.INJECT_ACCUMULATE)
    damp = [O.damping_profile(m, pml, float(np.float32(h)), O.DAMP_REFERENCE, 2000.0, float(np.float32(0.001)),
                              float(np.float32(3.4)), dtype=np.float64) for m in (nx, ny, nz)]
    d = tmp_path / "py"
    d.mkdir()
    print("[gpu] drop-in run 80x70x60, 41 steps, slices at step 40", flush=True)
    e = sgfd.ElasticWaveEquation(nt, None, pml, (nx, ny, nz), (dt, h, h, h), (b, lam, mu), (sx, sy, sz),
                                 (50, 40, 35), list(O.REF_COEF), f0, damping=damp)
    try:
        e.run(snap_every=20, snap_from=20, snap_dir=str(d))
        e.snapShootVolum(nt - 1, str(d))
    finally:
        e.close()
    w = O.ricker(f0, dt, nt, O.RICKER_CEIL_DELAY)
    _, F = O.run(p, b, lam, mu, w, None)
    vx, sxx, sxy = F[0], F[3], F[6]  # [iz][iy][ix]
    # (values, halo flags) in the writers' loop order
    halo = lambda n, i: i < 5 or i >= n - 5
    expect = {}
    for stem, f in (("velocityX", vx), ("stressXX", sxx), ("stressXY", sxy)):
        expect[f"{stem}SliceXX_YoZ_40.txt"] = (f, [(f[iz, iy, sx], halo(nz, iz) or halo(ny, iy))
                                                   for iz in range(nz) for iy in range(ny)])
        expect[f"{stem}SliceXX_XoZ_40.txt"] = (f, [(f[iz, sy, ix], halo(nz, iz) or halo(nx, ix))
                                                   for iz in range(nz) for ix in range(nx)])
        expect[f"{stem}SliceXX_XoY_40.txt"] = (f, [(f[sz, iy, ix], halo(ny, iy) or halo(nx, ix))
                                                   for iy in range(ny) for ix in range(nx)])
    expect["snapshot_0_40.txt"] = (sxx, [(sxx[iz, iy, ix], False) for iz in range(5, nz - 5)
                                         for iy in range(5, ny - 5) for ix in range(5, nx - 5)])
    checked = 0
    for name, (f, vh) in expect.items():
        got = (d / name).read_text().splitlines()
        want = _ref_text([v for v, _ in vh])
        hal = np.array([h for _, h in vh])
        assert len(got) == len(want), name
        g = np.array([float(x) for x in got])
        wv = np.array([float(x) for x in want])
        assert np.all(np.abs(g - wv) <= 1e-5 * np.abs(f).max()), name
        # the halo cells are never updated: "0" on exactly those lines of
        # the oracle's layout (pins the loop order against a transpose)
        assert all(got[i] == "0" for i in np.flatnonzero(hal)), name
        # every line is printf %g text of a float (6 significant digits)
        assert all(x == "%g" % float(x) for x in got), name
        # values above 1e-3 of the field maximum: at most one unit in the
        # 6th significant digit (printing) plus 1e-7 of the maximum (the
        # runs' own difference) apart
        big = np.abs(wv) >= 1e-3 * np.abs(f).max()
        unit = 10.0 ** (np.floor(np.log10(np.abs(wv[big]))) - 5)
        dev = np.abs(g[big] - wv[big]) - 1.01 * unit
        print(f"[gpu] {name}: worst {dev.max() / np.abs(f).max() if dev.size else 0:.2e} of the maximum beyond "
              "one printed unit", flush=True)
        assert np.all(dev <= 1e-7 * np.abs(f).max()), name
        checked += int(big.sum())
    assert checked > 1000, checked
