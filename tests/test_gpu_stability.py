"""Stability guard and FP64 energies on the GPU, through the C ABI.

* sg_set_medium refuses a model above the Courant limit (normalised Courant
  number 1.03) with SG_ERR_INVALID and the number in sg_last_error, accepts
  0.98, and accepts 1.03 once the limit is switched off.
* The energy check (every 10 steps) stops a run at Courant 1.03 with
  SG_ERR_UNSTABLE while every field is still finite: the run never reaches
  NaN.  A stable run with the same check runs to the end and gives the same
  bits as without it.
* sg_energy_full on the G1 golden run: against FP64 sums of the formula over
  the solver's own fields (the reduction itself, <= 1e-12) and over the
  compiled reference's fields of G1 (<= 1e-5); sg_energy returns the first
  two values; two calls give the same bits (fixed summation order).
"""
import sys
import numpy as np
import pytest

from conftest import load_golden, say

pytestmark = pytest.mark.gpu

ERR_INVALID, ERR_UNSTABLE = -1, -7


def _say(msg):
    say(msg)


@pytest.fixture(scope="module")
def sgfd():
    import sgfd as m

    m.lib()
    return m


def _case(sgfd, courant, n=48, pml=8, nt=300):
    """Homogeneous n^3 model and the dt that gives the normalised Courant number."""
    coef = sgfd.diff_coef("reference")
    cs = float(np.abs(coef.astype(np.float64)).sum())
    vp, vs, rho, h = 3000.0, 1732.0, 2000.0, 10.0
    dt = courant * h / (vp * cs * np.sqrt(3.0))
    shape = (n, n, n)
    mu = np.full(shape, rho * vs * vs, np.float32)
    lam = np.full(shape, rho * vp * vp, np.float32) - 2 * mu
    b = np.full(shape, 1.0 / rho, np.float32)
    w = sgfd.ricker(40.0, dt, nt, "none")
    return dict(n=n, pml=pml, nt=nt, dt=dt, h=h, coef=coef, b=b, lam=lam, mu=mu, w=w)


def _solver(sgfd, c, **stab):
    n = c["n"]
    s = sgfd.Solver(n, n, n, c["pml"], c["dt"], c["h"], c["h"], c["h"], coef=c["coef"], time_length=c["nt"])
    if stab:
        s.set_stability(**stab)
    s.set_damping(*[sgfd.damping_profile(n, c["pml"], c["h"], "quadratic", 3000.0)] * 3)
    return s


def test_set_medium_refuses_unstable_model(sgfd):
    _say("courant refusal 48^3")
    bad = _case(sgfd, 1.03)
    with _solver(sgfd, bad) as s:
        with pytest.raises(sgfd.SgError) as e:
            s.set_medium(bad["b"], bad["lam"], bad["mu"])
        assert e.value.code == ERR_INVALID and "Courant number 1.03" in str(e.value)
        assert s.courant() == 0.0  # nothing was taken
    good = _case(sgfd, 0.98)
    with _solver(sgfd, good) as s:
        s.set_medium(good["b"], good["lam"], good["mu"])
        assert s.courant() == pytest.approx(0.98, rel=1e-6)
    with _solver(sgfd, bad, courant_limit=0.0) as s:
        s.set_medium(bad["b"], bad["lam"], bad["mu"])
        assert s.courant() == pytest.approx(1.03, rel=1e-6)


def test_energy_check_stops_diverging_run(sgfd):
    _say("energy check stops a Courant 1.03 run, 48^3 P8")
    c = _case(sgfd, 1.03)
    with _solver(sgfd, c, courant_limit=0.0, check_interval=10, max_growth=2.0) as s:
        s.set_medium(c["b"], c["lam"], c["mu"])
        s.set_source(24, 24, 24, c["w"])
        with pytest.raises(sgfd.SgError) as e:
            s.step(c["nt"])
        assert e.value.code == ERR_UNSTABLE and "energy grew" in str(e.value)
        done = s.steps_done
        assert 10 <= done < c["nt"] and done % 10 == 0
        F = s.fields()
        assert np.all(np.isfinite(F))
        _say(f"stopped after step {done}; max |field| {np.abs(F).max():.3e}")


def test_energy_check_leaves_stable_run_unchanged(sgfd):
    _say("energy check on a stable run (Courant 0.98), 48^3 P8, 300 steps")
    c = _case(sgfd, 0.98)
    out = []
    for interval in (0, 10):
        with _solver(sgfd, c, check_interval=interval) as s:
            s.set_medium(c["b"], c["lam"], c["mu"])
            s.set_source(24, 24, 24, c["w"])
            s.set_stations(np.array([[30, 24, 24], [24, 30, 20]], np.int32), comps=("vx", "sxx"))
            s.run()
            out.append((s.stations(), s.fields()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def _energy_np(F, b, lam, mu):
    u = (slice(5, -5),) * 3
    v = F[0:3][(slice(None),) + u].astype(np.float64)
    sg = F[3:9][(slice(None),) + u].astype(np.float64)
    b, lam, mu = (a[u].astype(np.float64) for a in (b, lam, mu))
    ek = (0.5 * (v ** 2).sum(0) / b).sum()
    n2 = (sg[:3] ** 2).sum(0)
    sh2 = (sg[3:] ** 2).sum(0)
    tr = sg[:3].sum(0)
    es = (n2 + sh2).sum()
    ee = ((n2 + 2 * sh2) / (4 * mu) - lam * tr ** 2 / (4 * mu * (3 * lam + 2 * mu))).sum()
    return np.array([ek, es, ee])


def test_energy_full_against_fp64_sums_g1(oracle_mod, sgfd):
    O = oracle_mod
    meta, g = load_golden("g1_workload_hetero.npz")
    NX, NY, NZ = meta["total"]
    _say(f"energy_full on the G1 run {NX}x{NY}x{NZ}")
    b, lam, mu = g["b"], g["lam"], g["mu"]
    w = O.ricker(meta["f0"], meta["dt"], meta["nt"], O.RICKER_NO_DELAY)
    f = lambda x: float(np.float32(x))
    damp = [O.damping_profile(n, meta["pml"], f(h), O.DAMP_QUADRATIC, f(meta["damp_v"]), f(0.001), f(3.4),
                              dtype=np.float64) for n, h in ((NX, meta["dx"]), (NY, meta["dy"]), (NZ, meta["dz"]))]
    with sgfd.Solver(NX, NY, NZ, meta["pml"], meta["dt"], meta["dx"], meta["dy"], meta["dz"],
                     coef=O.EXACT_COEF, time_length=meta["nt"]) as s:
        s.set_medium(b, lam, mu)
        s.set_damping(*damp)
        s.set_source(*meta["src"], w)
        s.run()
        e = s.energy_full()
        e2 = s.energy_full()
        e_old = s.energy()
        F = s.fields()
    assert np.array_equal(e, e2)
    assert np.array_equal(e_old, e[:2])
    own = _energy_np(F, b, lam, mu)
    assert np.all(own > 0)
    assert np.max(np.abs(e - own) / own) < 1e-12
    ref = _energy_np(g["fields"], b, lam, mu)
    assert np.max(np.abs(e - ref) / ref) < 2e-5  # quadratic in fields that agree to 1e-5
