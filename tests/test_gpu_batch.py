"""Tile shapes, the autotuner and batched shots through the C ABI.

The per-point arithmetic does not depend on how points are grouped into
tiles, z-chunks or batches, so all of these must be bit-identical to a single
solver with the default tile."""
import sys
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _say(msg):
    say(msg)


@pytest.fixture(scope="module")
def sgfd_mod():
    import sgfd

    sgfd.lib()
    return sgfd


def _small_workload(sgfd, seed, nx, ny, nz, pml=6, nt=60, f0=20.0):
    from sgfd import configs

    g = sgfd.gaussian_field(nx, ny, nz, 4.0, seed, 4.0).astype(np.float64)
    vp = 3000.0 * (1 + 0.05 * g)
    b, lam, mu = configs.elastic_from_velocity(vp, vp / np.sqrt(3.0), 310.0 * vp ** 0.25)
    rng = np.random.default_rng(seed)
    lo = 5 + pml + 2
    src = tuple(int(rng.integers(lo, n - lo)) for n in (nx, ny, nz))
    st = np.array([[lo, lo, lo], [nx - lo, ny // 2, nz // 2], [src[0], src[1], 5 + pml]], np.int32)
    return configs.Workload(f"s{seed}", nx, ny, nz, pml, 1e-3, 10.0, 10.0, 10.0, nt, b, lam, mu, src,
                            sgfd.ricker(f0, 1e-3, nt, "none"), st, comps=("vx", "vy", "vz", "sxx"),
                            damp_v=float(vp.max()), vp_max=float(vp.max()))


def _run(wl, tile=None, graph=False):
    s = wl.solver()
    try:
        if tile:
            s.set_tile(*tile)
        if graph:
            s.set_graph(True)
        s.step(wl.nt)
        s.synchronize()
        return s.stations(), s.fields()
    finally:
        s.close()


def test_tile_shapes_are_bit_identical(sgfd_mod):
    wl = _small_workload(sgfd_mod, 3, 150, 77, 61)
    _say("tile sweep 150x77x61")
    tr0, F0 = _run(wl)
    for tile in ((4, 1, 4, 1, 0), (8, 2, 16, 2, 0), (16, 2, 8, 1, 7), (16, 1, 16, 1, 20), (4, 1, 8, 2, 51),
                 (8, 1, 4, 1, 13)):
        tr, F = _run(wl, tile)
        assert np.array_equal(F, F0), tile
        assert np.array_equal(tr, tr0), tile


def test_no_pml_lds_rows_are_bit_identical(sgfd_mod):
    """A grid without PML: with separate list launches its interior boxes
    load their staging rows straight into LDS (FP32); one stream runs the
    fused launch, which stages through registers.  Every box shape, ragged
    extents (150x77x61): the same bits."""
    wl = _small_workload(sgfd_mod, 9, 150, 77, 61, pml=0)
    _say("no PML, 150x77x61: streams 1 against 3, box shapes (4,1) (8,1) (8,2) (16,1) (16,2)")
    ref = None
    for tile in ((4, 1, 4, 1, 0), (8, 1, 4, 1, 0), (8, 2, 4, 1, 9), (16, 1, 4, 1, 20), (16, 2, 4, 1, 0)):
        for n in (1, 3):
            s = wl.solver()
            try:
                s.set_tile(*tile)
                s.set_streams(n)
                s.step(wl.nt)
                s.synchronize()
                out = (s.stations(), s.fields())
            finally:
                s.close()
            if ref is None:
                ref = out
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), (tile, n)
            assert np.isfinite(out[1]).all()


def test_stream_counts_are_bit_identical(sgfd_mod):
    wl = _small_workload(sgfd_mod, 5, 150, 77, 61)
    _say("streams 1..4, 150x77x61, eager and graph")
    ref = None
    for n in (1, 2, 3, 4):
        for graph in (False, True):
            s = wl.solver()
            try:
                s.set_streams(n)
                assert s.streams() == n
                s.set_graph(graph)
                s.step(wl.nt)
                s.synchronize()
                out = (s.stations(), s.fields())
            finally:
                s.close()
            if ref is None:
                ref = out
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), (n, graph)


def test_autotune_keeps_results(sgfd_mod):
    wl = _small_workload(sgfd_mod, 4, 96, 90, 70)
    _say("autotune 96x90x70")
    tr0, F0 = _run(wl)
    s = wl.solver()
    try:
        (ty, ny, pty, pny, zc, wv), ms = s.autotune(2)
        assert ty in (4, 8, 16) and pty in (4, 8, 16) and ny in (1, 2) and pny in (1, 2) and ms > 0
        assert wv in (0, 2, 3, 4)
        assert s.tile() == (ty, ny, pty, pny, zc, wv)
        assert s.streams() in (1, 2, 3) and (wv == 0 or s.streams() == 1)
        assert s.steps_done == 0
        s.step(wl.nt)
        s.synchronize()
        assert np.array_equal(s.fields(), F0)
        assert np.array_equal(s.stations(), tr0)
    finally:
        s.close()


@pytest.mark.parametrize("launch", [dict(ty=16, ny=1, pty=8, pny=1, zchunk=32, grid=0, graph=True),
                                    dict(ty=16, ny=2, pty=4, pny=1, zchunk=13, grid=64, graph=False),
                                    dict(ty=4, pty=4, zchunk=24, grid=0, graph=True, waves=0),
                                    dict(ty=4, pty=4, zchunk=16, grid=96, graph=True, waves=2),
                                    dict(ty=4, pty=4, zchunk=32, grid=0, graph=False, waves=3),
                                    dict(ty=4, pty=4, zchunk=9, grid=40, graph=True, waves=4)])
def test_batched_shots_match_single_shots(sgfd_mod, launch):
    sgfd = sgfd_mod
    dims = [(96, 80, 60), (130, 72, 64), (70, 101, 50), (64, 64, 80)]
    wls = [_small_workload(sgfd, 10 + i, *d) for i, d in enumerate(dims)]
    _say(f"batch of {len(wls)} shots, launch {launch}")
    ref = [_run(wl) for wl in wls]
    solvers = [wl.solver() for wl in wls]
    b = None
    try:
        b = sgfd.Batch(solvers)
        b.set_launch(**launch)
        b.step(wls[0].nt)
        b.synchronize()
        for s, (tr0, F0) in zip(solvers, ref):
            assert s.steps_done == wls[0].nt
            assert np.array_equal(s.fields(), F0)
            assert np.array_equal(s.stations(), tr0)
    finally:
        if b is not None:
            b.close()
        for s in solvers:
            s.close()


def _reconfigure(sgfd, s, wl, k):
    """Between steps: a new station set (another count and components), a
    new source point and wavelet, and the per-cell damping; each of these
    reallocates or adds device buffers behind the shot's descriptors."""
    st = np.array([[wl.nx // 2, wl.ny // 2, 5 + wl.pml + k], [wl.nx // 3, 12, 20]], np.int32)
    s.set_stations(st, comps=("vz", "sxx", "syz"))
    src = (wl.src[0] + 3, wl.src[1] - 2, wl.src[2] + 1)
    s.set_source(*src, -0.5 * sgfd.ricker(25.0, 1e-3, wl.nt, "none"))
    s.set_damping_local(*[sgfd.damping_profile(n, wl.pml, 10.0, "cuda", v=1.0, tau=4.3) for n in (wl.nx, wl.ny, wl.nz)])


@pytest.mark.parametrize("wait", [True, False])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_picks_up_shot_changes(sgfd_mod, graph, wait):
    """Shots reconfigured between batch steps (stations, source, damping)
    give the same results, bit for bit, as the same shots stepped alone with
    the same changes at the same step.  Without `wait` the setters are called
    while the batched steps may still run: each setter waits for the batch
    streams of the batches holding its solver, so nothing changes under a
    running sweep."""
    sgfd = sgfd_mod
    wls = [_small_workload(sgfd, 30 + i, *d) for i, d in enumerate([(96, 80, 60), (70, 101, 50)])]
    half = wls[0].nt // 2
    _say(f"batch of {len(wls)} shots reconfigured after {half} steps, graph={graph}")
    ref = []
    for k, wl in enumerate(wls):
        s = wl.solver()
        try:
            s.step(half)
            _reconfigure(sgfd, s, wl, k)
            s.step(wl.nt - half)
            ref.append((s.stations(), s.fields()))
        finally:
            s.close()
    solvers = [wl.solver() for wl in wls]
    b = None
    try:
        b = sgfd.Batch(solvers)
        b.set_launch(graph=graph)
        b.step(half)
        if wait:
            b.synchronize()
        for k, (s, wl) in enumerate(zip(solvers, wls)):
            _reconfigure(sgfd, s, wl, k)
        b.step(wls[0].nt - half)
        b.synchronize()
        for s, (tr0, F0) in zip(solvers, ref):
            assert np.array_equal(s.fields(), F0)
            assert np.array_equal(s.stations(), tr0)
    finally:
        if b is not None:
            b.close()
        for s in solvers:
            s.close()


def test_destroying_a_batched_solver_is_refused(sgfd_mod):
    """sg_destroy of a solver that a live batch holds returns SG_ERR_STATE
    and frees nothing (the batch would keep its device pointers); after the
    batch is destroyed the solver can be stepped and destroyed as usual."""
    sgfd = sgfd_mod
    wls = [_small_workload(sgfd, 40 + i, 64, 48, 40, nt=10) for i in range(2)]
    solvers = [wl.solver() for wl in wls]
    b = sgfd.Batch(solvers)
    try:
        with pytest.raises(sgfd.SgError, match="live batch"):
            solvers[0].close()
        b.step(5)
        b.synchronize()
    finally:
        b.close()
    try:
        solvers[0].step(5)
        assert solvers[0].steps_done == 10 and np.abs(solvers[0].fields()).max() > 0
    finally:
        for s in solvers:
            s.close()


def test_batched_accumulate_sources_match_single_shots(sgfd_mod):
    """The CUDA path's source in a batch: one shot with its source in the
    interior (kept in sigma), one inside the x PML (dropped after one
    velocity update); bit-identical to the same shots run alone."""
    sgfd = sgfd_mod
    wls = [_small_workload(sgfd, 50 + i, 96, 80, 60) for i in range(2)]
    wls[1].src = (8, 40, 30)  # x PML slab: columns 5..10
    ref = []
    for wl in wls:
        s = wl.solver(source_mode="accumulate")
        try:
            s.step(wl.nt)
            ref.append((s.stations(), s.fields()))
        finally:
            s.close()
    solvers = [wl.solver(source_mode="accumulate") for wl in wls]
    b = None
    try:
        b = sgfd.Batch(solvers)
        b.step(wls[0].nt)
        b.synchronize()
        for s, (tr0, F0) in zip(solvers, ref):
            assert np.array_equal(s.fields(), F0)
            assert np.array_equal(s.stations(), tr0)
    finally:
        if b is not None:
            b.close()
        for s in solvers:
            s.close()


@pytest.mark.parametrize("seed", [400, 401, 402])
def test_random_batches_match_single_shots(sgfd_mod, seed):
    """Three seeded random shots in one batch with another seeded geometry's
    box shapes, a persistent or a plain grid, graph or eager, stations
    decimated by 1..3: bit-identical to the shots run alone."""
    from conftest import random_case, random_workload

    sgfd = sgfd_mod
    rng = np.random.default_rng(seed)
    wls = [random_workload(seed * 10 + i, nt=30)[0] for i in range(3)]
    for wl in wls:
        wl.interval = int(rng.integers(1, 4))
    tile = random_case(seed)[2]
    grid, graph = int(rng.choice([0, 64, 300])), bool(seed % 2)
    _say(f"random batch {seed}: box shapes {tile}, grid {grid}, graph {graph}")
    ref = [_run(wl) for wl in wls]
    solvers = [wl.solver() for wl in wls]
    b = None
    try:
        b = sgfd.Batch(solvers)
        b.set_launch(*tile, grid=grid, graph=graph)
        b.step(30)
        b.synchronize()
        for s, (tr0, F0) in zip(solvers, ref):
            assert np.array_equal(s.fields(), F0)
            assert np.array_equal(s.stations(), tr0)
    finally:
        if b is not None:
            b.close()
        for s in solvers:
            s.close()


def test_failed_station_request_leaves_the_solver_unchanged(sgfd_mod):
    """A station set whose buffer does not fit the device (3000 stations x 9
    components x 2^26 steps, 7.2 TB) fails with SG_ERR_NOMEM, and the solver
    keeps its previous stations: it steps and records exactly like a solver
    that never saw the request."""
    sgfd = sgfd_mod
    wl = _small_workload(sgfd, 5, 64, 64, 64)
    st = np.array([[30, 30, 30], [34, 30, 30]], np.int32)

    def make(time_length):
        s = sgfd.Solver(64, 64, 64, wl.pml, wl.dt, wl.dx, wl.dy, wl.dz, coef=wl.coef, time_length=time_length)
        s.set_medium(wl.b, wl.lam, wl.mu)
        s.set_source(*wl.src, wl.wavelet)
        s.set_stations(st, comps=("vx", "sxx"))
        return s

    _say("failed station request, 64^3")
    with make(wl.nt) as r:
        r.step(wl.nt)
        ref = (r.stations(), r.fields())
    with make(1 << 26) as s:  # a 512 MiB station buffer
        with pytest.raises(sgfd.SgError) as e:
            s.set_stations(np.tile(st[:1], (3000, 1)), comps=sgfd.FIELDS)
        assert e.value.code == -3  # SG_ERR_NOMEM
        s.step(wl.nt)
        assert np.array_equal(s.stations(), ref[0])
        assert np.array_equal(s.fields(), ref[1])


def test_batch_autotune_keeps_results(sgfd_mod):
    """The batch autotuner (z-chunk, register bound, grid of the one-launch
    form, two per-list shapes) resets every shot; stepping afterwards gives
    the single shots' results bit for bit."""
    sgfd = sgfd_mod
    dims = [(96, 80, 60), (70, 101, 50), (64, 64, 80)]
    wls = [_small_workload(sgfd, 20 + i, *d) for i, d in enumerate(dims)]
    _say(f"batch autotune, {len(wls)} shots")
    ref = [_run(wl) for wl in wls]
    solvers = [wl.solver() for wl in wls]
    b = None
    try:
        b = sgfd.Batch(solvers)
        tile, grid, ms = b.autotune(2)
        _say(f"chose {tile} grid {grid}, {ms:.3f} ms per step")
        assert ms > 0 and tile[5] in (0, 2, 3, 4)
        assert all(s.steps_done == 0 for s in solvers)
        b.step(wls[0].nt)
        b.synchronize()
        for s, (tr0, F0) in zip(solvers, ref):
            assert np.array_equal(s.fields(), F0)
            assert np.array_equal(s.stations(), tr0)
    finally:
        if b is not None:
            b.close()
        for s in solvers:
            s.close()
