"""16-bit storage paths (FP16/BF16 with power-of-two field scales), their
rounding modes and the compensated time update, through the C ABI.

There is no 16-bit counterpart in the reference (its solvers are FP32), so
these results are checked against the FP64 build of the same kernels on the
same inputs; the tolerances are the measured error levels of each mode with
margin (plain FP16 rounding noise grows with the number of steps)."""
import sys
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _say(msg):
    say(msg)


@pytest.fixture(scope="module")
def setup():
    import sgfd
    from sgfd import configs

    sgfd.lib()
    nx, ny, nz, pml, nt = 96, 80, 72, 8, 240
    g = sgfd.gaussian_field(nx, ny, nz, 5.0, 11, 4.0).astype(np.float64)
    vp = 3000.0 * (1 + 0.05 * g)
    b, lam, mu = configs.elastic_from_velocity(vp, vp / 1.75, 310.0 * vp ** 0.25)
    st = np.array([[x, 40, 20] for x in range(20, 80, 6)], np.int32)
    wl = configs.Workload("p16", nx, ny, nz, pml, 1e-3, 10.0, 10.0, 10.0, nt, b, lam, mu, (48, 40, 36),
                          sgfd.ricker(18.0, 1e-3, nt, "none"), st, damp_v=float(vp.max()), vp_max=float(vp.max()),
                          f0=18.0)
    s = wl.solver("fp64")
    s.step(nt)
    ref = s.stations().astype(np.float64)
    s.close()
    return sgfd, wl, ref


def _run(wl, prec, rounding="nearest", comp=None, tile=None):
    s = wl.solver(prec)
    try:
        if tile:
            s.set_tile(*tile)
        s.set_rounding(rounding)
        if comp:
            s.set_compensation(comp)
        s.step(wl.nt)
        s.synchronize()
        return s.stations().astype(np.float64), s.fields()
    finally:
        s.close()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_pilot_scales_are_powers_of_two(setup):
    sgfd, wl, _ = setup
    sc = wl.pilot_scales()
    assert np.all(sc > 0)
    assert np.all(np.log2(sc) == np.round(np.log2(sc)))


def test_field_maxabs_matches_fields(setup):
    sgfd, wl, _ = setup
    s = wl.solver("fp32")
    try:
        s.step(120)
        m = s.field_maxabs()
        F = s.fields()
        assert np.array_equal(m, np.abs(F).reshape(9, -1).max(axis=1).astype(np.float32))
    finally:
        s.close()


@pytest.mark.parametrize("prec,tol", [("fp16", 2e-2), ("bf16", 1.2e-1)])
def test_16bit_storage_against_fp64(setup, prec, tol):
    sgfd, wl, ref = setup
    tr, F = _run(wl, prec)
    err = _rel(tr, ref)
    _say(f"{prec} nearest: seismogram rel-L2 {err:.3e}")
    assert np.isfinite(F).all() and err < tol


def test_compensation_reduces_fp16_error(setup):
    sgfd, wl, ref = setup
    plain = _rel(_run(wl, "fp16")[0], ref)
    comp = _rel(_run(wl, "fp16", comp="all")[0], ref)
    _say(f"fp16 plain {plain:.3e}, compensated {comp:.3e}")
    assert comp < 0.6 * plain and comp < 8e-3


def test_16bit_results_do_not_depend_on_tile_shape(setup):
    sgfd, wl, _ = setup
    for prec, rounding, comp in (("fp16", "nearest", None), ("fp16", "stochastic", None), ("fp16", "nearest", "all"),
                                 ("bf16", "stochastic", "velocity")):
        a = _run(wl, prec, rounding, comp)
        b = _run(wl, prec, rounding, comp, tile=(16, 1, 8, 1, 13))
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (prec, rounding, comp)


def test_stochastic_rounding_is_unbiased_enough(setup):
    sgfd, wl, ref = setup
    err = _rel(_run(wl, "fp16", "stochastic")[0], ref)
    _say(f"fp16 stochastic: {err:.3e}")
    assert err < 3e-2


def test_compensation_rejected_where_it_does_not_apply(setup):
    sgfd, wl, _ = setup
    s = wl.solver("fp32")
    try:
        with pytest.raises(sgfd.SgError):
            s.set_compensation("all")
    finally:
        s.close()
    s = wl.solver("fp16")
    try:
        s.step(2)
        with pytest.raises(sgfd.SgError):
            s.set_compensation("all")
        s.reset()
        s.set_compensation("all")
    finally:
        s.close()


@pytest.mark.parametrize("seed", [300, 301, 302, 303])
def test_16bit_random_geometries(seed):
    """Seeded random geometries and box shapes: FP16 and BF16 storage give
    the same bits with the default and the random box shape, and stay within
    their error levels of the FP64 build (fields rel-L2 2e-2 / 1.5e-1)."""
    import sgfd
    from conftest import random_workload

    sgfd.lib()
    wl, tile = random_workload(seed)
    s = wl.solver("fp64")
    try:
        s.step(wl.nt)
        F64 = s.fields(np.float64)
    finally:
        s.close()
    sc = wl.pilot_scales(steps=wl.nt)
    for prec, lim in (("fp16", 2e-2), ("bf16", 1.5e-1)):
        out = []
        for t in (None, tile):
            s = wl.solver(prec, scales=sc)
            try:
                if t:
                    s.set_tile(*t)
                s.step(wl.nt)
                out.append((s.stations(), s.fields(np.float64)))
            finally:
                s.close()
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]), prec
        err = float(np.linalg.norm(out[0][1] - F64) / np.linalg.norm(F64))
        _say(f"seed {seed} {prec}: fields rel-L2 {err:.2e} against FP64")
        assert np.isfinite(out[0][1]).all() and err < lim, prec
