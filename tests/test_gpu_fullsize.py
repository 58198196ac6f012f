"""Parity at the benchmark sizes, through the C ABI.

* 256^3 random medium (P16): the first 40 steps against the C oracle (its
  OpenMP build finishes them in seconds): every wavefield rel-L2 <= 1e-5.
* 256^3 random medium, the full 1000 steps: FP32 against the FP64 build on
  the 64x64 receivers, rel-L2 <= 1e-5, and on every wavefield, <= 3e-5
  (measured 6.7e-6 to 1.2e-5: the FP32 rounding of 1000 steps).
* 512x512x384 four-layer model (P20), the full 2000 steps: FP32 station
  traces and final wavefields against the FP64 build of the same kernels,
  rel-L2 <= 1e-5; FP16 and BF16 storage (scales from an FP32 pilot run)
  held below bounds above their measured levels (seismograms 7.6e-3 and
  6.2e-2).
* 1024x1024x640 four-layer model (P24), the full 1500 steps: FP32 and FP16
  storage against the FP64 build on the 256x256 surface receivers and on a
  seeded 2 % subsample of the points of vz and sxx, FP32 rel-L2 <= 1e-5.
* 1344x1344x1300 (2.35 G points, beyond 2^31 per array), homogeneous,
  source 100 cells below the top: after 20 steps the wave has spread at most
  200 cells, so a 440x440x330 grid with the source at the same distance from
  its top boundary must hold the same numbers, bit for bit (64-bit indexing
  of the large arrays, the source above linear index 2^31).
* 1800x1800x1300 with P24 (4.2 G points, 367 M split slots per split
  array, so each component's three splits together span more than 4 GiB in
  FP32): the same comparison against a 460x460x330 grid.  Every split array
  is addressed from its own base, and only one array has to stay below 4 GiB.
* 512x512x384, size-independent property: the update is linear in the
  source amplitude and scaling by 2 is exact in binary floating point, so a
  doubled wavelet must give doubled fields.  Where the leading edge of the
  wave falls through the subnormal range, rounding is absolute rather than
  relative and the two runs differ; measured, those differences climb about
  20 decades in 120 steps.  In FP64 (range floor 1e-308) every value above
  1e-200 must therefore be doubled bit for bit.  In FP32 velocities peak near
  1e-17, only 21 decades above the floor, so the differences reach the signal
  (rel-L2 ~1e-8) and FP32 is held to rel-L2 <= 1e-6.
"""
import sys
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _say(msg):
    say(msg)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def mods():
    import sgfd
    from sgfd import configs

    sgfd.lib()
    return sgfd, configs


def test_random256_first_steps_against_oracle(mods, oracle_mod):
    sgfd, configs = mods
    O = oracle_mod
    nt = 40
    wl = configs.random256(nt=nt)
    _say("random256 40 steps: GPU")
    s = wl.solver()
    try:
        s.step(nt)
        s.synchronize()
        F = s.fields()
    finally:
        s.close()
    _say("random256 40 steps: oracle")
    p = O.OracleParams(wl.nx, wl.ny, wl.nz, wl.pml, wl.dt, wl.dx, wl.dy, wl.dz, nt, wl.src,
                       coef=tuple(float(c) for c in wl.coef), damp_mode=O.DAMP_QUADRATIC, damp_v=wl.damp_v,
                       nthreads=16)
    _, Fo = O.run(p, wl.b, wl.lam, wl.mu, np.asarray(wl.wavelet[:nt], np.float32), None)
    for f in range(9):
        e = _rel(F[f], Fo[f])
        assert e < 1e-5, (sgfd.FIELDS[f], e)


def test_random256_fp32_against_fp64_full_run(mods):
    sgfd, configs = mods
    wl = configs.random256()
    res = {}
    for prec in ("fp64", "fp32"):
        _say(f"random256 {wl.nt} steps {prec}")
        s = wl.solver(prec)
        try:
            s.run()
            res[prec] = (s.stations().astype(np.float64), s.fields(np.float64))
        finally:
            s.close()
    assert res["fp64"][0].shape == (wl.nt, 64 * 64, 3)
    e_tr = _rel(res["fp32"][0], res["fp64"][0])
    _say(f"random256 receivers rel-L2 fp32 vs fp64: {e_tr:.3e}")
    assert e_tr < 1e-5
    # the final fields hold the FP32 rounding of 1000 steps of the scheme
    # itself (measured 6.7e-6 to 1.2e-5 per field), which any FP32
    # implementation carries; the parity bar against the FP32 oracle is the
    # first-steps test above
    for f in range(9):
        e = _rel(res["fp32"][1][f], res["fp64"][1][f])
        _say(f"random256 {sgfd.FIELDS[f]} rel-L2 fp32 vs fp64: {e:.3e}")
        assert e < 3e-5, sgfd.FIELDS[f]


def test_layered512_precisions_against_fp64_full_run(mods):
    sgfd, configs = mods
    wl = configs.layered512()
    ref_tr = ref_f = None
    err, ferr = {}, {}
    for prec in ("fp64", "fp32", "fp16", "bf16"):
        _say(f"layered512 2000 steps {prec}")
        s = wl.solver(prec)
        try:
            s.step(wl.nt)
            s.synchronize()
            tr = s.stations().astype(np.float64)
            if prec == "fp64":
                ref_tr, ref_f = tr, s.fields(np.float64)
                continue
            err[prec] = _rel(tr, ref_tr)
            # final wavefields, all nine together, against the FP64 run
            num = den = 0.0
            for f in range(9):
                d = s.field(f, np.float64)
                num += float(np.sum((d - ref_f[f]) ** 2))
                den += float(np.sum(ref_f[f] ** 2))
            ferr[prec] = float(np.sqrt(num / den))
        finally:
            s.close()
    _say("seismogram rel-L2 against fp64: " + ", ".join(f"{p} {e:.3e}" for p, e in err.items()))
    _say("final-field rel-L2 against fp64: " + ", ".join(f"{p} {e:.3e}" for p, e in ferr.items()))
    assert err["fp32"] < 1e-5 and ferr["fp32"] < 1e-5
    # 16-bit storage: held to its measured levels with margin (seismograms
    # 7.6e-3 / 6.2e-2 for FP16 / BF16); only upper bounds, so a more
    # accurate 16-bit mode passes
    # (measured: fields FP16 2.2e-2, BF16 1.8e-1)
    assert err["fp16"] < 1e-2 and ferr["fp16"] < 3e-2
    assert err["bf16"] < 8e-2 and ferr["bf16"] < 2.5e-1


def test_large1024_fp32_fp16_against_fp64_subsample(mods):
    sgfd, configs = mods
    _say("large1024: building the model")
    wl = configs.large1024()
    n = wl.nx * wl.ny * wl.nz
    idx = np.random.default_rng(1024).choice(n, size=n // 50, replace=False)
    res = {}
    for prec in ("fp64", "fp32", "fp16"):
        _say(f"large1024 1500 steps {prec}")
        s = wl.solver(prec)
        try:
            for k in range(0, wl.nt, 300):
                s.step(min(300, wl.nt - k))
                s.synchronize()
                _say(f"large1024 {prec}: {min(k + 300, wl.nt)} steps")
            res[prec] = [s.stations().astype(np.float64)] + [s.field(f, np.float64).ravel()[idx] for f in ("vz", "sxx")]
        finally:
            s.close()
    # FP32 at the parity bar; FP16 storage held to its measured level with
    # margin (upper bound only)
    # (FP16 measured: stations 6.7e-3, vz 9.7e-3, sxx 8.1e-3)
    for prec, tol in (("fp32", 1e-5), ("fp16", 2e-2)):
        for what, a, b in zip(("stations", "vz 2 %", "sxx 2 %"), res[prec], res["fp64"]):
            e = _rel(a, b)
            _say(f"large1024 {what} rel-L2 {prec} vs fp64: {e:.3e}")
            assert e < tol, (prec, what)


def test_layered512_linearity(mods):
    sgfd, configs = mods
    wl = configs.layered512(nt=120)
    out = {}
    for prec in ("fp64", "fp32"):
        for scale in (1.0, 2.0):
            _say(f"layered512 120 steps {prec}, wavelet x{scale:g}")
            s = wl.solver(prec)
            try:
                s.set_source(*wl.src, np.asarray(wl.wavelet, np.float32) * np.float32(scale))
                s.step(wl.nt)
                s.synchronize()
                out[prec, scale] = (s.field("vz", np.float64), s.field("sxy", np.float64))
            finally:
                s.close()
    # FP64: every value far above the subnormal range is exactly doubled
    for a, b in zip(out["fp64", 1.0], out["fp64", 2.0]):
        big = np.abs(b) > 1e-200
        assert big.sum() > 10_000_000
        assert np.array_equal(2 * a[big], b[big])
    # FP32: the same up to what passes through the FP32 subnormal range
    for a, b in zip(out["fp32", 1.0], out["fp32", 2.0]):
        e = _rel(2 * a, b)
        _say(f"fp32 doubled: rel-L2 {e:.2e}")
        assert e < 1e-6


def _split_slots(nx, ny, nz, p):
    """Slots of one split array: z slabs, y slabs of the z interior, x slabs
    of the y/z interior (the solver's split store, 64-element aligned)."""
    a = lambda n: (n + 63) // 64 * 64
    ux, uy, uz = nx - 10, ny - 10, nz - 10
    return (a(2 * p * ny * nx) + a((uz - 2 * p) * 2 * p * nx)
            + 2 * a((uz - 2 * p) * (uy - 2 * p) * p))


def _compare_with_small_grid(sgfd, pml, big_shape, big_src, small_shape, small_src):
    nt, h, dt = 20, 10.0, 1e-3
    rho, vp = 2000.0, 3000.0
    mu_v, lam_v = rho * (vp / np.sqrt(3)) ** 2, rho * vp ** 2 - 2 * rho * (vp / np.sqrt(3)) ** 2
    w = sgfd.ricker(15.0, dt, nt, "none")

    def run(nx, ny, nz, src):
        n = nx * ny * nz
        s = sgfd.Solver(nx, ny, nz, pml, dt, h, h, h, coef=sgfd.diff_coef("exact"), time_length=nt)
        try:
            b = np.full(n, 1 / rho, np.float32)
            lam = np.full(n, lam_v, np.float32)
            mu = np.full(n, mu_v, np.float32)
            s.set_medium(b, lam, mu)
            del b, lam, mu
            s.set_damping(*[sgfd.damping_profile(m, pml, h, "quadratic", vp) for m in (nx, ny, nz)])
            s.set_source(*src, w)
            s.step(nt)
            s.synchronize()
            return {f: s.field(f) for f in ("vz", "sxx", "sxy")}
        finally:
            s.close()

    _say("small grid %dx%dx%d" % small_shape)
    small = run(*small_shape, small_src)
    _say("large grid %dx%dx%d (%.2f G points)" % (*big_shape, np.prod(big_shape) / 1e9))
    big = run(*big_shape, big_src)
    lo = [big_src[a] - small_src[a] for a in range(3)]
    sl = (slice(lo[2], lo[2] + small_shape[2]), slice(lo[1], lo[1] + small_shape[1]), slice(lo[0], lo[0] + small_shape[0]))
    for f in small:
        assert np.abs(small[f]).max() > 0
        assert np.array_equal(big[f][sl], small[f]), f
        # nothing outside the compared block was reached
        assert np.count_nonzero(big[f]) == np.count_nonzero(small[f]), f
        del big[f]


def test_grid_beyond_2p31_points_matches_small_grid(mods):
    sgfd, _ = mods
    big_shape, big_src = (1344, 1344, 1300), (672, 672, 1200)
    assert np.prod(big_shape) > 2 ** 31 and (big_src[2] * big_shape[1] + big_src[1]) * big_shape[0] > 2 ** 31
    _compare_with_small_grid(sgfd, 10, big_shape, big_src, (440, 440, 330), (220, 220, 230))


def test_split_store_beyond_4gib_per_component_matches_small_grid(mods):
    sgfd, _ = mods
    big_shape, pml = (1800, 1800, 1300), 24
    slots = _split_slots(*big_shape, pml)
    assert 3 * slots * 4 > 2 ** 32 > slots * 4
    _compare_with_small_grid(sgfd, pml, big_shape, (900, 900, 1200), (460, 460, 330), (230, 230, 230))
