"""Stability limit of the scheme, on the CPU.

* ``sg_courant_number`` (what ``sg_set_medium`` checks before it accepts a
  model) against its closed form dt * max(vp) * sum|c_n| *
  sqrt(1/dx^2 + 1/dy^2 + 1/dz^2), vp = sqrt((lambda + 2 mu) b), for
  homogeneous and heterogeneous models and unequal spacings.
* The limit 1 is the scheme's real stability boundary: the C oracle (a
  restatement of the reference CPU solver, no PML) run at normalised Courant
  numbers 0.98 and 1.03 stays bounded in the first case and leaves the float
  range or grows by many decades in the second.  A wrong factor (sqrt(3),
  the coefficient sum, 2 in the symbol) would move the boundary by 30 % or
  more and fail one of the two runs.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def sgfd():
    import sgfd as m

    m.lib()
    return m


def _homog(n, vp, vs, rho):
    mu = np.full(n, rho * vs * vs, np.float32)
    lam = np.full(n, rho * vp * vp, np.float32) - 2 * mu
    b = np.full(n, 1.0 / rho, np.float32)
    return b, lam, mu


def test_courant_number_closed_form(sgfd):
    coef = sgfd.diff_coef("reference")
    cs = float(np.abs(coef.astype(np.float64)).sum())
    for (dx, dy, dz), dt in [((10.0, 10.0, 10.0), 1e-3), ((5.0, 7.5, 12.5), 4e-4), ((12.5, 12.5, 12.5), 1e-3)]:
        b, lam, mu = _homog(64, 3000.0, 1732.0, 2000.0)
        vp = np.sqrt((lam.astype(np.float64) + 2 * mu) * b)
        want = dt * vp.max() * cs * np.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2)
        got = sgfd.courant_number(b, lam, mu, dt, dx, dy, dz, coef)
        assert got == pytest.approx(want, rel=1e-6)
    # equal spacing: the familiar 10th-order bound vp dt / h <= 0.4385
    b, lam, mu = _homog(8, 4385.0, 2500.0, 2000.0)
    assert sgfd.courant_number(b, lam, mu, 1e-3, 10.0, 10.0, 10.0, coef) == pytest.approx(1.0, abs=2e-3)


def test_courant_number_takes_the_fastest_point(sgfd):
    rng = np.random.default_rng(3)
    vp = 3000.0 * (1 + 0.1 * rng.standard_normal(5000))
    rho = 310.0 * vp ** 0.25
    vs = vp / 1.75
    b, lam, mu = sgfd.elastic_from_velocity(vp, vs, rho)
    coef = sgfd.diff_coef("exact")
    vpx = np.sqrt((lam.astype(np.float64) + 2 * mu.astype(np.float64)) * b.astype(np.float64))
    dt = float(np.float32(8e-4))  # the library takes dt as a float
    want = dt * vpx.max() * np.abs(coef.astype(np.float64)).sum() * np.sqrt(3) / 10.0
    assert sgfd.courant_number(b, lam, mu, 8e-4, 10.0, 10.0, 10.0, coef) == pytest.approx(want, rel=1e-9)


@pytest.mark.parametrize("bad", ["mu", "b", "bulk", "nan"])
def test_courant_number_rejects_bad_models(sgfd, bad):
    b, lam, mu = _homog(16, 3000.0, 1732.0, 2000.0)
    if bad == "mu":
        mu[3] = 0.0
    elif bad == "b":
        b[5] = -1.0
    elif bad == "bulk":
        lam[7] = -mu[7]  # 3 lambda + 2 mu < 0
    else:
        lam[2] = np.nan
    with pytest.raises(sgfd.SgError) as e:
        sgfd.courant_number(b, lam, mu, 1e-3, 10.0, 10.0, 10.0, sgfd.diff_coef("reference"))
    assert e.value.code == -1


def _oracle_growth(O, courant, nt=300):
    """Largest |field| of a 36^3, P0 oracle run at the given normalised
    Courant number (homogeneous, centred Ricker source)."""
    n = 36
    coef = O.REF_COEF
    cs = sum(abs(c) for c in coef)
    vp, vs, rho, h = 3000.0, 1732.0, 2000.0, 10.0
    dt = courant * h / (vp * cs * np.sqrt(3.0))
    shape = (n, n, n)
    mu = np.full(shape, rho * vs * vs, np.float32)
    lam = np.full(shape, rho * vp * vp, np.float32) - 2 * mu
    b = np.full(shape, 1.0 / rho, np.float32)
    w = O.ricker(40.0, dt, nt, O.RICKER_NO_DELAY)
    p = O.OracleParams(n, n, n, 0, dt, h, h, h, nt, (n // 2, n // 2, n // 2), coef=coef)
    _, F = O.run(p, b, lam, mu, w, None)
    with np.errstate(invalid="ignore", over="ignore"):
        return float(np.nanmax(np.abs(F))) if np.all(np.isfinite(F)) else np.inf


def test_limit_is_the_stability_boundary(oracle_mod):
    stable = _oracle_growth(oracle_mod, 0.98)
    unstable = _oracle_growth(oracle_mod, 1.03)
    assert np.isfinite(stable) and stable < 1e-3
    assert not np.isfinite(unstable) or unstable > 1e6 * stable
