"""sg_energy_noalloc through the C ABI: the energy reduction with its work
array in the library's fixed device scratch instead of a solver-allocated
buffer.  Two calls in a row, the second on a smaller grid (so the scratch
still holds the larger grid's partial sums), each bit-identical to
sg_energy on the same state; a grid with more updated z planes than the
scratch holds is refused with SG_ERR_UNSUPPORTED while sg_energy still
serves it."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DP = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def sgfd():
    import sgfd as m

    m.lib()
    return m


def _solver(sgfd, nx, ny, nz, pml, seed):
    rng = np.random.default_rng(seed)
    vp = 3000.0 * (1 + 0.05 * rng.standard_normal((nz, ny, nx)))
    b, lam, mu = sgfd.elastic_from_velocity(vp, vp / 1.8, np.full_like(vp, 2000.0))
    s = sgfd.Solver(nx, ny, nz, pml, 1e-3, 10.0, 10.0, 10.0, coef=sgfd.diff_coef("exact"), time_length=200)
    s.set_medium(b, lam, mu)
    s.set_damping(*[sgfd.damping_profile(n, pml, 10.0, "quadratic", 3000.0) for n in (nx, ny, nz)])
    s.set_source(nx // 2, ny // 2, nz // 2, sgfd.ricker(20.0, 1e-3, 200, "none"))
    return s


def test_energy_noalloc_matches_energy_twice_second_smaller(sgfd):
    L = sgfd.lib()
    for (nx, ny, nz, pml, nt) in ((96, 80, 72, 8, 60), (40, 36, 30, 4, 40)):
        print(f"[gpu] energy_noalloc {nx}x{ny}x{nz}, {nt} steps", flush=True)
        with _solver(sgfd, nx, ny, nz, pml, nx) as s:
            s.step(nt)
            a = np.full(2, np.nan)
            e = np.full(2, np.nan)
            assert L.sg_energy_noalloc(s.handle, a.ctypes.data_as(DP)) == 0
            assert L.sg_energy(s.handle, e.ctypes.data_as(DP)) == 0
            assert np.all(a > 0)
            assert np.array_equal(a, e), (a, e)


def test_energy_noalloc_refuses_more_planes_than_the_scratch(sgfd):
    L = sgfd.lib()
    print("[gpu] energy_noalloc on 11x11x16395 (16385 updated planes)", flush=True)
    with _solver(sgfd, 11, 11, 16395, 0, 1) as s:
        s.step(2)
        a = np.zeros(2)
        assert L.sg_energy_noalloc(s.handle, a.ctypes.data_as(DP)) == -6  # SG_ERR_UNSUPPORTED
        assert b"16384" in L.sg_last_error()
        assert L.sg_energy(s.handle, a.ctypes.data_as(DP)) == 0
        assert np.all(np.isfinite(a))
