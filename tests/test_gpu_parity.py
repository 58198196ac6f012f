"""Parity of the HIP solver (through the C ABI) with the CPU oracle.

Tolerance: relative L2 <= 1e-5 for FP32 wavefields and station traces after
the full run (the stated parity bar); FP64 against the FP64 oracle <= 1e-11.
Both kernel variants (tiled, simple) are checked.
"""
import sys
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, random_case as _random_case, say

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-11


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _say(msg):
    say(msg)


@pytest.fixture(scope="module")
def sgfd_mod():
    import sgfd

    sgfd.lib()
    return sgfd


def _solve(sgfd, p, b, lam, mu, wavelet, stations, comps, kernel="tiled", precision="fp32", damp=None,
           source_mode="transient", graph=False, tile=None, streams=None):
    s = sgfd.Solver(p.nx, p.ny, p.nz, p.pml, p.dt, p.dx, p.dy, p.dz, coef=p.coef, time_length=p.nt,
                    precision=precision, kernel=kernel, source_mode=source_mode)
    try:
        s.set_medium(b, lam, mu)
        if damp is not None:
            s.set_damping(*damp)
        s.set_source(*p.src, wavelet)
        s.set_stations(stations, comps=comps)
        if graph:
            s.set_graph(True)
        if tile:
            s.set_tile(*tile)
        if streams:
            s.set_streams(streams)
        s.run()
        return s.stations(), s.fields(np.float64 if precision == "fp64" else np.float32)
    finally:
        s.close()


def _damp_for(O, p):
    # the oracle holds V, R, tau and the spacings as float; use the same values
    f = lambda x: float(np.float32(x))
    return [O.damping_profile(n, p.pml, f(h), p.damp_mode, f(p.damp_v), f(p.damp_r), f(p.damp_tau), dtype=np.float64)
            for n, h in ((p.nx, p.dx), (p.ny, p.dy), (p.nz, p.dz))]


def _golden_case(O, name):
    meta, g = load_golden(name)
    NX, NY, NZ = meta["total"]
    native = meta["kind"] == "native"
    p = O.OracleParams(NX, NY, NZ, meta["pml"], meta["dt"], meta["dx"], meta["dy"], meta["dz"], meta["nt"],
                       tuple(meta["src"]), coef=O.REF_COEF if native else O.EXACT_COEF,
                       damp_mode=O.DAMP_REFERENCE if native else O.DAMP_QUADRATIC, damp_v=meta["damp_v"])
    if native:
        rho = np.full((NZ, NY, NX), 2500.0, np.float32)
        V2 = np.float32(2000.0 * 2000.0)
        b = (np.float32(1.0) / rho).astype(np.float32)
        lam = (rho * V2).astype(np.float32)
        mu = (rho * V2 / np.float32(3.0)).astype(np.float32)
        w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_CEIL_DELAY)
    else:
        b, lam, mu = g["b"], g["lam"], g["mu"]
        w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_NO_DELAY)
    return meta, g, p, b, lam, mu, w


@pytest.mark.parametrize("kernel", ["tiled", "simple"])
@pytest.mark.parametrize("name", ["g0_native_homog.npz", "g1_workload_hetero.npz"])
def test_golden_fp32(oracle_mod, sgfd_mod, name, kernel):
    O = oracle_mod
    meta, g, p, b, lam, mu, w = _golden_case(O, name)
    _say(f"golden {name} kernel={kernel} grid={meta['total']} nt={meta['nt']}")
    tr, F = _solve(sgfd_mod, p, b, lam, mu, w, g["stations"], sgfd_mod.FIELDS, kernel=kernel, damp=_damp_for(O, p))
    # golden fields come from the compiled reference itself
    assert _rel(F, g["fields"]) < TOL32
    assert _rel(tr, g["traces"]) < TOL32
    for f in range(9):
        assert not F[f][:5].any() and not F[f][-5:].any()
        assert not F[f][:, :5].any() and not F[f][:, -5:].any()
        assert not F[f][:, :, :5].any() and not F[f][:, :, -5:].any()


# The tiled FP64 sweeps took GPU memory faults on this case and on the 100^3
# FP64 case below; no mistake has been found in their code (every address
# they form lies inside the arrays), so sg_create refuses them and FP64 runs
# the per-point kernels.  The tiled cases stay skipped.
FP64_TILED_SKIP = "tiled FP64 sweeps disabled after GPU memory faults of unknown cause"


@pytest.mark.parametrize("kernel", [pytest.param("tiled", marks=pytest.mark.skip(reason=FP64_TILED_SKIP)),
                                    "simple"])
def test_golden_fp64_matches_fp64_oracle(oracle_mod, sgfd_mod, kernel):
    O = oracle_mod
    meta, g, p, b, lam, mu, w = _golden_case(O, "g1_workload_hetero.npz")
    _say(f"fp64 g1 kernel={kernel}")
    tr, F = _solve(sgfd_mod, p, b, lam, mu, w, g["stations"], sgfd_mod.FIELDS, kernel=kernel, precision="fp64",
                   damp=_damp_for(O, p))
    p64 = p
    tr64, F64 = O.run(p64, b.astype(np.float64), lam.astype(np.float64), mu.astype(np.float64),
                      w.astype(np.float64), g["stations"], dtype=np.float64)
    assert _rel(F, F64) < TOL64
    # the GPU station buffer is handed out as float32
    assert _rel(tr, tr64) < 1e-6


def _plain100(O):
    meta, g = load_golden("g2_plain100_traces.npz")
    NX, NY, NZ = meta["total"]
    rho = np.full((NZ, NY, NX), meta["rho"], np.float64)
    mu = (rho * meta["vs"] ** 2)
    lam = (rho * meta["vp"] ** 2 - 2 * mu).astype(np.float32)
    mu = mu.astype(np.float32)
    b = (1 / rho).astype(np.float32)
    p = O.OracleParams(NX, NY, NZ, meta["pml"], meta["dt"], meta["dx"], meta["dy"], meta["dz"], meta["nt"],
                       tuple(meta["src"]), coef=O.EXACT_COEF, damp_mode=O.DAMP_QUADRATIC, damp_v=meta["damp_v"])
    w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_NO_DELAY)
    return meta, g, p, b, lam, mu, w


_ORACLE_CACHE = {}


def _oracle_plain100(O, p, b, lam, mu, w, nt, dtype):
    key = (nt, np.dtype(dtype).name)
    if key not in _ORACLE_CACHE:
        import dataclasses

        q = dataclasses.replace(p, nt=nt)
        if dtype == np.float64:
            _ORACLE_CACHE[key] = O.run(q, b.astype(np.float64), lam.astype(np.float64), mu.astype(np.float64),
                                       w.astype(np.float64), None, dtype=np.float64)[1]
        else:
            _ORACLE_CACHE[key] = O.run(q, b, lam, mu, w, None)[1]
    return _ORACLE_CACHE[key]


@pytest.mark.parametrize("kernel", ["tiled", "simple"])
def test_plain100_full_run(oracle_mod, sgfd_mod, kernel):
    """100^3 incl. a 10-cell PML, 500 steps.

    * station traces of the whole run against the compiled reference's golden
      vectors: rel-L2 <= 1e-5;
    * all nine wavefields at step 150 (wave inside the grid) against the
      oracle: rel-L2 <= 1e-5;
    * the final wavefields (step 500) are the residual left after the wave has
      gone through the PML (~1e-3 of the peak); there the reference path's own
      FP32 result is 3.9e-4 from the FP64 oracle, so this solver's FP32 result
      must be at least as close to FP64 as the reference's (factor 1.5).
    """
    O = oracle_mod
    meta, g, p, b, lam, mu, w = _plain100(O)
    _say(f"plain100 100^3 500 steps kernel={kernel}")
    s = sgfd_mod.Solver(p.nx, p.ny, p.nz, p.pml, p.dt, p.dx, p.dy, p.dz, coef=p.coef, time_length=p.nt,
                        kernel=kernel)
    try:
        s.set_medium(b, lam, mu)
        s.set_damping(*_damp_for(O, p))
        s.set_source(*p.src, w)
        s.set_stations(g["stations"], comps=("vx", "vy", "vz", "sxx"))
        s.step(150)
        F150 = s.fields()
        s.step(p.nt - 150)
        s.synchronize()
        tr, F = s.stations(), s.fields()
    finally:
        s.close()
    # velocity traces jointly (vy is 20-40x smaller than vx/vz on this line,
    # which lies next to the source's y-symmetry plane), then each dominant one
    assert _rel(tr[:, :, :3], g["traces"][:, :, :3]) < TOL32
    for c in (0, 2, 3):
        assert _rel(tr[:, :, c], g["traces"][:, :, c]) < TOL32
    _say("plain100 oracle runs (150 steps f32, 500 steps f32/f64)")
    O150 = _oracle_plain100(O, p, b, lam, mu, w, 150, np.float32)
    for f in range(9):
        assert _rel(F150[f], O150[f]) < TOL32, sgfd_mod.FIELDS[f]
    O500 = _oracle_plain100(O, p, b, lam, mu, w, p.nt, np.float32)
    O500d = _oracle_plain100(O, p, b, lam, mu, w, p.nt, np.float64)
    ref_err = _rel(O500, O500d)
    assert ref_err < 1e-3
    assert _rel(F, O500d) < 1.5 * ref_err
    assert _rel(F[:, 2::5, 2::5, 2::5], g["fields_sub"]) < 2 * ref_err


@pytest.mark.parametrize("kernel", [pytest.param("tiled", marks=pytest.mark.skip(reason=FP64_TILED_SKIP)),
                                    "simple"])
def test_plain100_fp64_matches_fp64_oracle(oracle_mod, sgfd_mod, kernel):
    O = oracle_mod
    meta, g, p, b, lam, mu, w = _plain100(O)
    _say(f"plain100 fp64 kernel={kernel}")
    tr, F = _solve(sgfd_mod, p, b, lam, mu, w, g["stations"], ("vx", "vy", "vz", "sxx"), precision="fp64",
                   damp=_damp_for(O, p), kernel=kernel)
    O500d = _oracle_plain100(O, p, b, lam, mu, w, p.nt, np.float64)
    assert _rel(F, O500d) < 1e-9


def test_reference_cuda_station_file(sgfd_mod):
    """The reference's CUDA station file (..._concise/stationData.txt) before
    any PML interaction, with the CUDA path's accumulating source."""
    sgfd = sgfd_mod
    ref = np.loadtxt(os.path.join(GOLDEN, "stationData_concise_gpu.txt"))
    _say("reference CUDA station file, 300x200x100, 260 steps")
    nx, ny, nz, pml, nt = 300, 200, 100, 10, 260
    grid = sgfd.GRID(nx, ny, nz)
    delta = sgfd.Delta(5e-4, 10.0, 10.0, 10.0)
    rho = np.float32(2600.0)
    lam0 = np.float32(2600.0 * 5000.0 ** 2)
    shape = (nz, ny, nx)
    med = (np.full(shape, np.float32(1.0) / rho, np.float32), np.full(shape, lam0, np.float32),
           np.full(shape, np.float32(lam0 / np.float32(3.0)), np.float32))
    ewq = sgfd.ElasticWaveEquation(nt, None, pml, grid, delta, med, (150, 100, 50), (160, 110, 60),
                                   sgfd.diff_coef("reference"), 20.0)
    try:
        data = ewq.run()[:, 0, 0].astype(np.float64)
    finally:
        ewq.close()
    n = 250
    assert np.argmax(data[:n]) == np.argmax(ref[:n]) == 235
    assert _rel(data[:n], ref[:n]) < 1e-4



def test_propagate_per_step_matches_run(sgfd_mod):
    """The reference's per-step entry, propagate(timeIndex, sourceValue)
    (..._concise/ElasticWaveEquation.cu:194-206), driven with
    sourceFunction(it) for every step, gives the same bits as run(); a
    time index out of order is refused."""
    sgfd = sgfd_mod
    nx, ny, nz, pml, nt = 64, 56, 48, 6, 80
    shape = (nz, ny, nx)
    med = (np.full(shape, np.float32(1 / 2600.0), np.float32), np.full(shape, np.float32(2600.0 * 4.0e6), np.float32),
           np.full(shape, np.float32(2600.0 * 4.0e6 / 3.0), np.float32))
    args = (nt, None, pml, (nx, ny, nz), (5e-4, 10.0, 10.0, 10.0), med, (32, 28, 24), [(40, 28, 24), (20, 30, 30)],
            sgfd.diff_coef("reference"), 25.0)
    a = sgfd.ElasticWaveEquation(*args)
    try:
        ref = a.run(snap_every=0)
        Fa = a.solver.fields()
    finally:
        a.close()
    b = sgfd.ElasticWaveEquation(*args)
    try:
        with pytest.raises(ValueError):
            b.propagate(1, 0.0)
        for it in range(nt):
            b.propagate(it, b.sourceFunction(it))
        assert np.array_equal(b.solver.stations(), ref) and np.array_equal(b.solver.fields(), Fa)
        assert np.abs(ref).max() > 0
    finally:
        b.close()

def test_reference_cuda_station_file_all_rows(sgfd_mod):
    """All 1500 rows of the reference's CUDA station file, with the CUDA
    path's accumulating source AND its per-cell damping (tau 4.3, local S
    velocity): the rows after ~500 are the residue the PML leaves (1e-3 of
    the peak) and pin the absorbing layer itself.  Measured: 1.2e-5 over the
    whole file, <= 5e-4 within each late window."""
    sgfd = sgfd_mod
    ref = np.loadtxt(os.path.join(GOLDEN, "stationData_concise_gpu.txt"))
    nx, ny, nz, pml, nt = 300, 200, 100, 10, ref.shape[0]
    _say(f"reference CUDA station file, CUDA damping, {nt} steps")
    rho = np.float32(2600.0)
    lam0 = np.float32(2600.0 * 5000.0 ** 2)
    shape = (nz, ny, nx)
    med = (np.full(shape, np.float32(1.0) / rho, np.float32), np.full(shape, lam0, np.float32),
           np.full(shape, np.float32(lam0 / np.float32(3.0)), np.float32))
    ewq = sgfd.ElasticWaveEquation(nt, None, pml, (nx, ny, nz), (5e-4, 10.0, 10.0, 10.0), med, (150, 100, 50),
                                   (160, 110, 60), sgfd.diff_coef("reference"), 20.0, damping="cuda")
    try:
        data = ewq.run(snap_every=0)[:, 0, 0].astype(np.float64)
    finally:
        ewq.close()
    assert _rel(data, ref) < 5e-5
    for a, b in ((250, 500), (500, 800), (800, 1100), (1100, nt)):
        assert _rel(data[a:b], ref[a:b]) < 2e-3, (a, b)


@pytest.mark.parametrize("kernel,streams", [("tiled", None), ("tiled", 3), ("simple", None)])
def test_no_pml_against_oracle(oracle_mod, sgfd_mod, kernel, streams):
    """P = 0 (the reference's no-PML variants: only the 5-cell halo around
    the updated range): heterogeneous 70x66x62, 120 steps, all fields
    rel-L2 <= 1e-5 against the oracle.  With separate list launches (three
    streams) the interior boxes of a grid without PML load their staging rows
    straight into LDS; one stream runs the single fused launch."""
    O = oracle_mod
    nx, ny, nz, nt = 70, 66, 62, 120
    rng = np.random.default_rng(5)
    vp = 3000.0 * (1 + 0.05 * rng.standard_normal((nz, ny, nx)))
    rho = 2000.0 + 100.0 * rng.standard_normal((nz, ny, nx))
    mu = (rho * (vp / 1.8) ** 2).astype(np.float32)
    lam = (rho * vp ** 2 - 2 * rho * (vp / 1.8) ** 2).astype(np.float32)
    b = (1 / rho).astype(np.float32)
    p = O.OracleParams(nx, ny, nz, 0, 8e-4, 10.0, 10.0, 10.0, nt, (35, 30, 31), coef=O.EXACT_COEF,
                       damp_mode=O.DAMP_QUADRATIC, damp_v=3000.0)
    w = O.ricker(20.0, 8e-4, nt, O.RICKER_NO_DELAY)
    st = np.array([[20, 20, 20], [50, 40, 30]], np.int32)
    _say(f"no PML, 70x66x62, {kernel}, streams {streams}")
    tr, F = _solve(sgfd_mod, p, b, lam, mu, w, st, sgfd_mod.FIELDS, kernel=kernel, damp=_damp_for(O, p),
                   streams=streams)
    trO, FO = O.run(p, b, lam, mu, w, st)
    for f in range(9):
        assert _rel(F[f], FO[f]) < TOL32, sgfd_mod.FIELDS[f]
    assert _rel(tr, trO) < TOL32



# (nx, ny, nz, pml, source): one updated point; a grid that is all PML
# (2*pml = n - 10, no interior box); x one column past two and one tiles;
# fewer x columns than a tile; y all PML with two rows; thin z
EXTREME = [(11, 11, 11, 0, (5, 5, 5)), (30, 30, 30, 10, (12, 15, 17)), (129, 17, 14, 2, (100, 8, 6)),
           (65, 23, 19, 3, (40, 11, 9)), (37, 90, 21, 4, (18, 60, 10)), (75, 12, 80, 1, (40, 6, 50))]


@pytest.mark.parametrize("kernel", ["tiled", "simple"])
@pytest.mark.parametrize("geom", EXTREME, ids=lambda g: "x".join(map(str, g[:3])) + f"p{g[3]}")
def test_extreme_shapes_against_oracle(oracle_mod, sgfd_mod, geom, kernel):
    """Degenerate and ragged grids (the extremes sg_create accepts), a
    heterogeneous medium, 40 steps: every field and the station traces
    rel-L2 <= 1e-5 against the oracle, and the halo stays exactly zero."""
    O = oracle_mod
    nx, ny, nz, pml, src = geom
    nt = 40
    rng = np.random.default_rng(nx * ny + nz)
    vp = 3000.0 * (1 + 0.05 * rng.standard_normal((nz, ny, nx)))
    rho = 2000.0 + 100.0 * rng.standard_normal((nz, ny, nx))
    mu = (rho * (vp / 1.8) ** 2).astype(np.float32)
    lam = (rho * vp ** 2 - 2 * rho * (vp / 1.8) ** 2).astype(np.float32)
    b = (1 / rho).astype(np.float32)
    p = O.OracleParams(nx, ny, nz, pml, 8e-4, 10.0, 10.0, 10.0, nt, src, coef=O.REF_COEF,
                       damp_mode=O.DAMP_REFERENCE, damp_v=3000.0)
    w = O.ricker(20.0, 8e-4, nt, O.RICKER_NO_DELAY)
    st = np.array([src, [5, 5, 5], [nx - 6, ny - 6, nz - 6]], np.int32)
    _say(f"extreme {nx}x{ny}x{nz} P{pml} {kernel}")
    tr, F = _solve(sgfd_mod, p, b, lam, mu, w, st, sgfd_mod.FIELDS, kernel=kernel, damp=_damp_for(O, p))
    trO, FO = O.run(p, b, lam, mu, w, st)
    assert np.abs(FO).max() > 0
    for f in range(9):
        assert _rel(F[f], FO[f]) < TOL32, sgfd_mod.FIELDS[f]
        assert not F[f][:5].any() and not F[f][-5:].any()
        assert not F[f][:, :5].any() and not F[f][:, -5:].any()
        assert not F[f][:, :, :5].any() and not F[f][:, :, -5:].any()
    assert _rel(tr, trO) < TOL32


def test_medium_from_velocity_matches_medium(sgfd_mod):
    """sg_set_medium_velocity (Vp, Vs, rho) gives the same run, bit for bit,
    as sg_set_medium with the moduli the library forms from them."""
    sgfd = sgfd_mod
    nx, ny, nz, pml, nt = 60, 50, 40, 4, 30
    rng = np.random.default_rng(11)
    vp = (3000 * (1 + 0.05 * rng.standard_normal((nz, ny, nx)))).astype(np.float32)
    vs = (vp / np.float32(1.75)).astype(np.float32)
    rho = (310 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    out = []
    for use_velocity in (True, False):
        s = sgfd.Solver(nx, ny, nz, pml, 8e-4, 10.0, 10.0, 10.0, time_length=nt)
        try:
            if use_velocity:
                s.set_medium_velocity(vp, vs, rho)
            else:
                s.set_medium(*sgfd.elastic_from_velocity(vp, vs, rho))
            s.set_source(30, 25, 20, sgfd.ricker(20.0, 8e-4, nt, "none"))
            s.step(nt)
            out.append(s.fields())
        finally:
            s.close()
    assert np.abs(out[0]).max() > 0 and np.array_equal(out[0], out[1])


RANDOM_SEEDS = range(24)


@pytest.mark.parametrize("seed", RANDOM_SEEDS)
def test_random_geometries_against_oracle(oracle_mod, sgfd_mod, seed):
    """Random extents, PML widths, box shapes and z-chunks (seeded), a
    heterogeneous medium, 30 steps: the tiled sweeps against the oracle,
    every field and the traces rel-L2 <= 1e-5."""
    O = oracle_mod
    (nx, ny, nz), pml, tile, src, st = _random_case(seed)
    nt = 30
    rng = np.random.default_rng(100 + seed)
    vp = 3000.0 * (1 + 0.05 * rng.standard_normal((nz, ny, nx)))
    rho = 2000.0 + 100.0 * rng.standard_normal((nz, ny, nx))
    mu = (rho * (vp / 1.8) ** 2).astype(np.float32)
    lam = (rho * vp ** 2 - 2 * rho * (vp / 1.8) ** 2).astype(np.float32)
    b = (1 / rho).astype(np.float32)
    p = O.OracleParams(nx, ny, nz, pml, 8e-4, 10.0, 10.0, 10.0, nt, src, coef=O.EXACT_COEF,
                       damp_mode=O.DAMP_QUADRATIC, damp_v=3000.0)
    w = O.ricker(20.0, 8e-4, nt, O.RICKER_NO_DELAY)
    _say(f"random case {seed}: {nx}x{ny}x{nz} P{pml} tile {tile} source {src}")
    tr, F = _solve(sgfd_mod, p, b, lam, mu, w, st, sgfd_mod.FIELDS, damp=_damp_for(O, p), tile=tile)
    trO, FO = O.run(p, b, lam, mu, w, st)
    for f in range(9):
        assert _rel(F[f], FO[f]) < TOL32, sgfd_mod.FIELDS[f]
    assert _rel(tr, trO) < TOL32


@pytest.mark.parametrize("kernel", ["tiled", "simple"])
@pytest.mark.parametrize("where", ["interior", "pml"])
def test_accumulate_source_against_oracle(oracle_mod, sgfd_mod, where, kernel):
    """The CUDA path's source (SG_SOURCE_ACCUMULATE) against the oracle's
    restatement of it, 60 steps, a station on the source point: at an
    interior source the injections stay in sigma; inside the PML (x slab)
    each is dropped after one velocity update, as the CUDA path's split
    rebuild does.  Every field and the traces rel-L2 <= 1e-5."""
    O = oracle_mod
    nx, ny, nz, pml, nt = 70, 60, 50, 8, 60
    src = (35, 30, 25) if where == "interior" else (9, 30, 25)
    rng = np.random.default_rng(9)
    vp = 3000.0 * (1 + 0.05 * rng.standard_normal((nz, ny, nx)))
    rho = 2000.0 + 100.0 * rng.standard_normal((nz, ny, nx))
    mu = (rho * (vp / 1.8) ** 2).astype(np.float32)
    lam = (rho * vp ** 2 - 2 * rho * (vp / 1.8) ** 2).astype(np.float32)
    b = (1 / rho).astype(np.float32)
    p = O.OracleParams(nx, ny, nz, pml, 8e-4, 10.0, 10.0, 10.0, nt, src, coef=O.EXACT_COEF,
                       damp_mode=O.DAMP_QUADRATIC, damp_v=3000.0, inject_mode=O.INJECT_ACCUMULATE)
    w = O.ricker(20.0, 8e-4, nt, O.RICKER_NO_DELAY)
    st = np.array([src, [src[0] + 3, src[1], src[2]], [src[0], src[1] - 4, src[2] + 2]], np.int32)
    _say(f"accumulate source, {where}, {kernel}")
    tr, F = _solve(sgfd_mod, p, b, lam, mu, w, st, sgfd_mod.FIELDS, kernel=kernel, damp=_damp_for(O, p),
                   source_mode="accumulate")
    trO, FO = O.run(p, b, lam, mu, w, st)
    for f in range(9):
        assert _rel(F[f], FO[f]) < TOL32, sgfd_mod.FIELDS[f]
    assert _rel(tr, trO) < TOL32
    # inside the PML this is the transient float injection; in the interior
    # the kept injections make a different run
    p.inject_mode = O.INJECT_FLOAT
    trT, _ = O.run(p, b, lam, mu, w, st, fields=False)
    assert (_rel(tr, trT) < TOL32) == (where == "pml")


@pytest.mark.parametrize("seed", [500, 501, 502])
def test_local_damping_random_geometries(sgfd_mod, seed):
    """The CUDA path's per-cell damping (sg_set_damping_local) on seeded
    random geometries, both source modes: the tiled sweeps give the same
    bits under two box shapes and agree with the per-point kernels
    (rel-L2 <= 1e-5)."""
    from conftest import random_workload

    sgfd = sgfd_mod
    wl, tile = random_workload(seed)
    g = [sgfd.damping_profile(n, wl.pml, 10.0, "cuda", v=1.0, tau=4.3) for n in (wl.nx, wl.ny, wl.nz)]
    for mode in ("transient", "accumulate"):
        out = []
        for kern, t in (("tiled", None), ("tiled", tile), ("simple", None)):
            s = wl.solver(kernel=kern, source_mode=mode)
            try:
                s.set_damping_local(*g)
                if t:
                    s.set_tile(*t)
                s.step(wl.nt)
                out.append((s.stations(), s.fields()))
            finally:
                s.close()
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]), mode
        assert _rel(out[0][1], out[2][1]) < TOL32 and _rel(out[0][0], out[2][0]) < TOL32, mode

def test_graph_replay_matches_eager(oracle_mod, sgfd_mod):
    O = oracle_mod
    meta, g, p, b, lam, mu, w = _golden_case(O, "g1_workload_hetero.npz")
    _say("graph replay vs eager, g1")
    d = _damp_for(O, p)
    tr1, F1 = _solve(sgfd_mod, p, b, lam, mu, w, g["stations"], sgfd_mod.FIELDS, damp=d)
    tr2, F2 = _solve(sgfd_mod, p, b, lam, mu, w, g["stations"], sgfd_mod.FIELDS, damp=d, graph=True)
    assert np.array_equal(F1, F2)
    assert np.array_equal(tr1, tr2)


def test_energy_decays_after_source(oracle_mod, sgfd_mod):
    sgfd = sgfd_mod
    n, pml = 60, 10
    _say("energy decay 60^3 pml 10")
    rho = np.full((n, n, n), 2000.0)
    mu = (rho * 1732.0 ** 2).astype(np.float32)
    lam = (rho * 3000.0 ** 2).astype(np.float32) - 2 * mu
    b = (1 / rho).astype(np.float32)
    with sgfd.Solver(n, n, n, pml, 1e-3, 10.0, 10.0, 10.0, coef=sgfd.diff_coef("exact"), time_length=600) as s:
        s.set_medium(b, lam, mu)
        s.set_damping(*[sgfd.damping_profile(n, pml, 10.0, "quadratic", 3000.0)] * 3)
        s.set_source(n // 2, n // 2, n // 2, sgfd.ricker(20.0, 1e-3, 600, "none"))
        s.step(150)
        e1 = s.energy()
        s.step(450)
        e2 = s.energy()
    assert e1[0] > 0 and e2[0] < 0.05 * e1[0]
    assert np.isfinite(e2).all()


def test_step_timed_and_bytes(oracle_mod, sgfd_mod):
    sgfd = sgfd_mod
    _say("step_timed 64^3")
    with sgfd.Solver(64, 64, 64, 8, 1e-3, 10.0, 10.0, 10.0, time_length=20) as s:
        n = 64
        s.set_medium(np.full(n ** 3, 1 / 2000.0, np.float32), np.full(n ** 3, 6e9, np.float32),
                     np.full(n ** 3, 6e9, np.float32))
        s.set_source(32, 32, 32, sgfd.ricker(15.0, 1e-3, 20))
        v, st, tot = s.step_timed(20)
        assert v > 0 and st > 0 and tot >= v + st
        vb, sb = s.sweep_bytes()
        u, P = 54, 8
        interior = (u - 2 * P) ** 3
        assert vb > 13 * 4 * u ** 3 and sb > 17 * 4 * u ** 3
        assert vb < 28 * 4 * u ** 3
        assert vb - 13 * 4 * interior > 0


def test_invalid_inputs_raise(sgfd_mod):
    sgfd = sgfd_mod
    with sgfd.Solver(32, 32, 32, 4, 1e-3, 10.0, 10.0, 10.0, time_length=4) as s:
        with pytest.raises(sgfd.SgError):
            s.set_source(2, 16, 16, np.ones(4, np.float32))  # inside the halo
        with pytest.raises(sgfd.SgError):
            s.set_stations([[40, 0, 0]])
        bad = np.ones(32 ** 3, np.float32)
        bad[100] = -1
        with pytest.raises(sgfd.SgError):
            s.set_medium(bad, bad, bad)


def test_cpp_dropin_program_reproduces_reference_station_file(sgfd_mod, tmp_path):
    """The reference's CUDA main program, ported to the C++ drop-in header
    (with the CUDA path's damping), run for all 1500 steps: its
    stationData.txt matches the reference's own file, and it writes the
    reference's slice snapshots (every 100 steps from 500)."""
    import subprocess

    from test_abi import _build_example

    exe = _build_example(tmp_path)
    _say("C++ drop-in program, 1500 steps")
    out = subprocess.run([str(exe), "1500"], cwd=tmp_path, check=True, capture_output=True, text=True, timeout=300)
    assert "recorded 1500 samples" in out.stdout
    got = np.loadtxt(tmp_path / "stationData.txt")
    ref = np.loadtxt(os.path.join(GOLDEN, "stationData_concise_gpu.txt"))
    assert _rel(got, ref) < 5e-5
    assert (tmp_path / "stressXXSliceXX_XoY_1400.txt").exists()
    assert len(list(tmp_path.glob("*SliceXX_*.txt"))) == 9 * 10
