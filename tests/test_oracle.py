"""Pins the CPU oracle (oracle/sg_oracle.c) before it is trusted as the checker.

* G0/G1/G2 are golden vectors written by the UNMODIFIED reference CPU solver
  (tests/golden/make_golden.py); the oracle must reproduce them bit for bit.
* stationData_concise_gpu.txt is the one numeric file the reference ships
  (3DGPUwithPML_1DArraySingleGPU_concise/stationData.txt, written by its CUDA
  solver).  Its rows before any PML interaction pin the interior stencil, the
  source (with the CUDA variant's accumulating injection), the time order and
  the recording.
* Closed forms and invariants pin what those leave open.
"""
import math

import numpy as np
import pytest

from conftest import load_golden


def _model_homog(shape, vp, vs, rho):
    rho = np.full(shape, rho, np.float64)
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    return (1 / rho).astype(np.float32), lam.astype(np.float32), mu.astype(np.float32)


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------- closed forms

def test_exact_coefficients_satisfy_taylor_conditions(oracle_mod):
    c = np.array(oracle_mod.EXACT_COEF, np.float64)
    odd = np.array([2 * n - 1 for n in range(1, 6)], np.float64)
    # first derivative: sum c_n (2n-1) = 1; odd moments 3,5,7,9 vanish
    assert abs(np.sum(c * odd) - 1.0) < 1e-15
    for m in (3, 5, 7, 9):
        assert abs(np.sum(c * odd ** m)) < 1e-9 * np.sum(np.abs(c) * odd ** m)
    # the reference literals are these rounded to 6 significant digits
    r = np.array(oracle_mod.REF_COEF, np.float64)
    assert np.all(np.abs(r - c) <= 5e-6 * np.abs(c) + 1e-9)


def test_ricker_matches_closed_form(oracle_mod):
    f0, dt, nt = 15.0, 1e-3, 300
    for mode, t0 in ((oracle_mod.RICKER_NO_DELAY, 0.0),
                     (oracle_mod.RICKER_CEIL_DELAY, math.ceil(1.0 / (np.float32(f0) * np.float32(dt))))):
        s = oracle_mod.ricker(f0, dt, nt, mode)
        t = (np.arange(nt) - t0) * dt - 1.0 / f0
        a = (np.pi * f0 * t) ** 2
        exact = (1 - 2 * a) * np.exp(-a)
        assert np.max(np.abs(s - exact)) < 2e-6
        assert int(np.argmax(s)) == int(round(t0 + 1.0 / (f0 * dt)))


def test_damping_closed_form(oracle_mod):
    n, pml, h = 40, 8, 10.0
    d = oracle_mod.damping_profile(n, pml, h, oracle_mod.DAMP_QUADRATIC, 3000.0, 1e-3, dtype=np.float64)
    d0 = -3 * 3000.0 * math.log(1e-3) / (2 * pml * h)
    for i in range(n):
        if 5 <= i < 5 + pml:
            dist = 5 + pml - i
        elif n - 5 - pml <= i < n - 5:
            dist = i - (n - 5 - pml) + 1
        else:
            assert d[i] == 0.0
            continue
        assert d[i] == pytest.approx(d0 * (dist / pml) ** 2, rel=1e-6)
    # the reference profile is symmetric: outermost updated cell has dist = pml
    assert d[5] == d[n - 6] and d[5 + pml - 1] == d[n - 5 - pml]
    r = oracle_mod.damping_profile(n, pml, h, oracle_mod.DAMP_REFERENCE, 2000.0, 1e-3, 3.4, dtype=np.float64)
    assert r[5] == pytest.approx(3.0 * 3.4 * 2000.0 / (pml * h), rel=1e-12)


# ------------------------------------------------------------- golden vectors

def _params_from_meta(O, meta, native):
    NX, NY, NZ = meta["total"]
    if native:
        return O.OracleParams(NX, NY, NZ, meta["pml"], meta["dt"], meta["dx"], meta["dy"], meta["dz"],
                              meta["nt"], tuple(meta["src"]), coef=O.REF_COEF, damp_mode=O.DAMP_REFERENCE,
                              damp_v=meta["damp_v"])
    return O.OracleParams(NX, NY, NZ, meta["pml"], meta["dt"], meta["dx"], meta["dy"], meta["dz"],
                          meta["nt"], tuple(meta["src"]), coef=O.EXACT_COEF, damp_mode=O.DAMP_QUADRATIC,
                          damp_v=meta["damp_v"])


def test_g0_native_constants_bitexact(oracle_mod):
    O = oracle_mod
    meta, g = load_golden("g0_native_homog.npz")
    NX, NY, NZ = meta["total"]
    rho = np.full((NZ, NY, NX), 2500.0, np.float32)
    V2 = np.float32(2000.0 * 2000.0)
    b = (np.float32(1.0) / rho).astype(np.float32)
    lam = (rho * V2).astype(np.float32)
    mu = (rho * V2 / np.float32(3.0)).astype(np.float32)
    p = _params_from_meta(O, meta, native=True)
    w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_CEIL_DELAY)
    assert np.array_equal(w, g["wavelet"])
    tr, F = O.run(p, b, lam, mu, w, g["stations"])
    assert np.array_equal(tr, g["traces"])
    assert np.array_equal(F, g["fields"])
    assert np.abs(F).max() > 0
    for a, key in enumerate(("damp_x", "damp_y", "damp_z")):
        d = O.damping_profile(meta["total"][a], meta["pml"], (meta["dx"], meta["dy"], meta["dz"])[a],
                              O.DAMP_REFERENCE, meta["damp_v"])
        assert np.array_equal(d, g[key])


def test_g1_heterogeneous_bitexact(oracle_mod):
    O = oracle_mod
    meta, g = load_golden("g1_workload_hetero.npz")
    p = _params_from_meta(O, meta, native=False)
    w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_NO_DELAY)
    assert np.array_equal(w, g["wavelet"])
    tr, F = O.run(p, g["b"], g["lam"], g["mu"], w, g["stations"])
    assert np.array_equal(tr, g["traces"])
    assert np.array_equal(F, g["fields"])
    for a, key in enumerate(("damp_x", "damp_y", "damp_z")):
        d = O.damping_profile(meta["total"][a], meta["pml"], (meta["dx"], meta["dy"], meta["dz"])[a],
                              O.DAMP_QUADRATIC, meta["damp_v"])
        assert np.array_equal(d, g[key])


def test_g2_plain100_traces_bitexact(oracle_mod):
    O = oracle_mod
    meta, g = load_golden("g2_plain100_traces.npz")
    NX, NY, NZ = meta["total"]
    b, lam, mu = _model_homog((NZ, NY, NX), meta["vp"], meta["vs"], meta["rho"])
    p = _params_from_meta(O, meta, native=False)
    w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_NO_DELAY)
    tr, F = O.run(p, b, lam, mu, w, g["stations"], comp_mask=0b1111)
    assert np.array_equal(tr, g["traces"])
    assert np.array_equal(F[:, 2::5, 2::5, 2::5], g["fields_sub"])
    for f in range(9):
        assert np.sum(F[f].astype(np.float64) ** 2) == pytest.approx(meta["field_sumsq"][f], rel=1e-12)


def test_reference_gpu_station_file(oracle_mod):
    """Rows 0..249 of the reference CUDA run.

    Reference setup (..._concise/main.cu:76-120): 300x200x100 total, pml 10,
    dt 0.5 ms, dx 10 m, f0 20 Hz, rho 2600, lambda = rho*5000^2, mu = lambda/3,
    source grid/2, station grid/2 + pml, sxx recorded.  The nearest PML face is
    35 cells from the source (z), so a 100^3 box with the same source-station
    offset (10,10,10) records the same signal until the first PML interaction
    (after row 250).  The CUDA solver keeps the injected stress in its unsplit
    interior, hence ORACLE_INJECT_ACCUMULATE; the file holds 6 significant
    digits of an FMA-contracted unsplit evaluation, hence a tolerance.
    """
    import os
    from conftest import GOLDEN

    O = oracle_mod
    ref = np.loadtxt(os.path.join(GOLDEN, "stationData_concise_gpu.txt"), dtype=np.float64)
    assert ref.shape == (1500,)
    nt, N = 250, 100
    rho = np.float32(2600.0)
    lam0 = np.float32(2600.0 * 5000.0 ** 2)
    shape = (N, N, N)
    b = np.full(shape, np.float32(1.0) / rho, np.float32)
    lam = np.full(shape, lam0, np.float32)
    mu = np.full(shape, np.float32(lam0 / np.float32(3.0)), np.float32)
    p = O.OracleParams(N, N, N, 10, 5e-4, 10.0, 10.0, 10.0, nt, (50, 50, 50), coef=O.REF_COEF,
                       inject_mode=O.INJECT_ACCUMULATE)
    w = O.ricker(20.0, 5e-4, nt, O.RICKER_CEIL_DELAY)
    tr, _ = O.run(p, b, lam, mu, w, [[60, 60, 60]], comp_mask=1 << 3, fields=False)
    got = tr[:, 0, 0].astype(np.float64)
    assert np.argmax(got) == np.argmax(ref[:nt]) == 235
    assert _rel_l2(got, ref[:nt]) < 1e-4
    # the CPU path's transient injection gives a different (un-integrated) signal
    p.inject_mode = O.INJECT_FLOAT
    tr2, _ = O.run(p, b, lam, mu, w, [[60, 60, 60]], comp_mask=1 << 3, fields=False)
    assert _rel_l2(tr2[:, 0, 0], ref[:nt]) > 0.5


# ------------------------------------------------------------------ invariants

def _small_case(O, n=30, pml=4, nt=80, f0=25.0):
    b, lam, mu = _model_homog((n, n, n), 3000.0, 1732.0, 2000.0)
    c = n // 2
    p = O.OracleParams(n, n, n, pml, 1e-3, 10.0, 10.0, 10.0, nt, (c, c, c), coef=O.EXACT_COEF,
                       damp_mode=O.DAMP_QUADRATIC, damp_v=3000.0)
    w = O.ricker(f0, 1e-3, nt, O.RICKER_NO_DELAY)
    return p, b, lam, mu, w


def test_linearity_in_source_amplitude_is_exact(oracle_mod):
    O = oracle_mod
    p, b, lam, mu, w = _small_case(O)
    _, F1 = O.run(p, b, lam, mu, w)
    _, F2 = O.run(p, b, lam, mu, (2 * w).astype(np.float32))
    assert np.array_equal(F2, 2 * F1)


def test_halo_stays_zero_and_axis_symmetry(oracle_mod):
    O = oracle_mod
    p, b, lam, mu, w = _small_case(O)
    _, F = O.run(p, b, lam, mu, w)
    n, c = p.nx, p.nx // 2
    for f in range(9):
        a = F[f]
        assert not a[:5].any() and not a[-5:].any()
        assert not a[:, :5].any() and not a[:, -5:].any()
        assert not a[:, :, :5].any() and not a[:, :, -5:].any()
    sxx, syy, szz = F[3], F[4], F[5]
    line_x = sxx[c, c, c + 1:n - 5]
    line_y = syy[c, c + 1:n - 5, c]
    line_z = szz[c + 1:n - 5, c, c]
    scale = np.abs(line_x).max()
    assert scale > 0
    # x<->y maps every split sum onto itself with its terms swapped, so the
    # float results are identical; x<->z re-associates (a+b)+c, so only
    # rounding-level agreement is expected.
    assert np.array_equal(line_x, line_y)
    assert np.max(np.abs(line_x - line_z)) < 1e-4 * scale


def test_fp64_oracle_agrees_with_fp32(oracle_mod):
    O = oracle_mod
    meta, g = load_golden("g1_workload_hetero.npz")
    p = _params_from_meta(O, meta, native=False)
    w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_NO_DELAY)
    tr32, F32 = O.run(p, g["b"], g["lam"], g["mu"], w, g["stations"])
    tr64, F64 = O.run(p, g["b"].astype(np.float64), g["lam"].astype(np.float64), g["mu"].astype(np.float64),
                      w.astype(np.float64), g["stations"], dtype=np.float64)
    assert _rel_l2(F32, F64) < 1e-5
    assert _rel_l2(tr32, tr64) < 1e-5


@pytest.mark.parametrize("where", ["interior", "pml"])
def test_accumulate_injection_closed_form(oracle_mod, where):
    """With zero buoyancy the velocities stay 0, so sigma at the source is
    the injection alone.  ACCUMULATE (the CUDA path): at an interior point
    sxx = syy = szz = the running float sum of dt*dt*s(it) after every step
    (the unsplit interior keeps each injection); inside the PML the split
    rebuild drops it, so 0.  The CPU form (DOUBLE) drops it everywhere."""
    import numpy as np

    O = oracle_mod
    n, pml, nt, dt = 24, 3, 12, 1e-3
    src = (12, 12, 12) if where == "interior" else (6, 12, 12)
    z = np.zeros((n, n, n), np.float32)
    lam = np.full((n, n, n), 1e9, np.float32)
    w = O.ricker(20.0, dt, nt, O.RICKER_NO_DELAY)
    st = np.array([src], np.int32)
    p = O.OracleParams(n, n, n, pml, dt, 10.0, 10.0, 10.0, nt, src, inject_mode=O.INJECT_ACCUMULATE)
    tr, F = O.run(p, z, lam, lam, w, st)
    acc = np.float32(0)
    exp = []
    d2 = np.float32(np.float32(dt) * np.float32(dt))
    for it in range(nt):
        acc = np.float32(acc + np.float32(d2 * w[it]))
        exp.append(acc if where == "interior" else np.float32(0))
    for c in (3, 4, 5):
        assert np.array_equal(tr[:, 0, c], np.array(exp, np.float32))
    assert not F[:3].any()
    p.inject_mode = O.INJECT_DOUBLE
    tr, F = O.run(p, z, lam, lam, w, st)
    assert not tr.any() and not F.any()
