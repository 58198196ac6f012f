"""CPU checks of the random-medium oracle (oracle/medium.py) and of the raw
float32 model files, pinned by published values, closed forms, brute force
and statistics rather than by restating the oracle's own formulas."""
import os

import numpy as np
import pytest

import medium as M


def test_splitmix64_published_outputs():
    # the first three outputs of splitmix64 seeded with 0 (state advanced by
    # the golden gamma 0x9e3779b97f4a7c15 before each mix)
    g = 0x9E3779B97F4A7C15
    states = np.array([0, g, (2 * g) % 2**64], np.uint64)
    assert [int(v) for v in M.splitmix64(states)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_phases_are_uniform_and_seeded():
    th = M.phases(200_000, 7)
    assert th.min() >= 0 and th.max() < 2 * np.pi
    # mean and variance of U[0, 2pi): pi and pi^2 / 3
    assert abs(th.mean() - np.pi) < 0.02 and abs(th.var() - np.pi**2 / 3) < 0.03
    assert np.array_equal(th, M.phases(200_000, 7)) and not np.array_equal(th, M.phases(200_000, 8))


def test_signed_index_is_numpy_fftfreq():
    for n in (1, 2, 5, 6, 9, 300):
        assert np.array_equal(M.signed_index(n), np.rint(np.fft.fftfreq(n) * n))


def test_spectrum_closed_forms():
    # reference form: (a b c) at k = 0; one axis at k = 1 (s = l)
    P = M.spectrum(11, 8, 6, M.KIND_REFERENCE, a=(2.0, 3.0, 4.0), order=1.0, l=5.0)
    assert P[0, 0, 0] == 24.0
    assert np.isclose(P[0, 0, 5], 24.0 / (1 + 4.0))
    assert np.isclose(P[0, 2, 0], 24.0 / (1 + (0.4 * 0.4 * 9.0)))
    # Gaussian: exp(-k^2 a^2 / 4), k = 2 pi s / n
    G = M.spectrum(8, 8, 8, M.KIND_GAUSSIAN, a=(2.0, 1.0, 1.0))
    assert np.isclose(G[0, 0, 2], np.exp(-(np.pi / 2) ** 2 * 4 / 4))
    assert np.isclose(G[0, 0, 6], G[0, 0, 2])  # s = -2
    # von Karman: (1 + k^2 a^2)^-(order + 3/2)
    V = M.spectrum(8, 8, 8, M.KIND_VON_KARMAN, a=(1.0, 3.0, 1.0), order=0.5)
    assert np.isclose(V[0, 1, 0], (1 + (np.pi / 4) ** 2 * 9) ** -2.0)


def test_inverse_transform_matches_brute_force_dft():
    nx, ny, nz = 5, 4, 3
    for kind in (M.KIND_REFERENCE, M.KIND_GAUSSIAN):
        psi = M.psi_grid(nx, ny, nz, kind, a=(1.5, 1.0, 2.0), order=1.0, l=2.0, seed=3)
        f = M.brute_inverse_dft(psi)
        want = np.abs(f) if kind == M.KIND_REFERENCE else f.real
        np.testing.assert_allclose(M.raw_field(nx, ny, nz, kind, a=(1.5, 1.0, 2.0), order=1.0, l=2.0, seed=3), want,
                                   rtol=1e-10, atol=1e-10)


def test_centred_spectrum_gives_the_same_modulus():
    # the script centres its spectrum on bin L/2; |ifftn| of the shifted
    # spectrum equals |ifftn| of the unshifted one
    psi = M.psi_grid(10, 8, 6, M.KIND_REFERENCE, a=(3.0, 3.0, 3.0), order=2.0, l=4.0, seed=1)
    a = np.abs(np.fft.ifftn(psi))
    b = np.abs(np.fft.ifftn(np.fft.fftshift(psi)))
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)


def test_normalisation_invariants():
    d = M.random_medium(20, 18, 16, "reference", a=(3.0, 3.0, 3.0), order=2.0, l=6.0, seed=5, v0=0.0, eps=0.1)
    assert abs(d.mean()) < 1e-15
    assert np.isclose(np.sqrt(np.mean(d * d)), 0.1, rtol=1e-12)
    v = M.random_medium(20, 18, 16, "reference", a=(3.0, 3.0, 3.0), order=2.0, l=6.0, seed=5, v0=2000.0, eps=0.1)
    np.testing.assert_allclose(v, 2000.0 * (1 + d), rtol=1e-14)


@pytest.mark.slow
def test_gaussian_medium_correlation():
    a = 4.0
    f = M.random_medium(64, 64, 64, "gaussian", a=(a, a, a), seed=11, v0=0.0, eps=1.0)
    lags = np.arange(7)
    np.testing.assert_allclose(M.sample_correlation_x(f, lags), np.exp(-(lags / a) ** 2), atol=0.05)


@pytest.mark.slow
def test_von_karman_correlation_follows_its_spectrum():
    # Re of a random-phase field has correlation sum_k P(k) cos(k r) / sum_k P(k)
    # (Wiener-Khinchin on the periodic grid); the spectrum's form is pinned by
    # the closed-form test above.  (Order 1/2 is exp(-r/a) in the continuum;
    # on a grid its k^-4 tail beyond Nyquist is cut, which raises the small-lag
    # correlation, so the grid's own expectation is the pin.)
    a, n = 3.0, 64
    f = M.random_medium(n, n, n, "von_karman", a=(a, a, a), order=0.5, seed=12, v0=0.0, eps=1.0)
    P = M.spectrum(n, n, n, M.KIND_VON_KARMAN, a=(a, a, a), order=0.5)
    lags = np.arange(7)
    kx = 2 * np.pi * M.signed_index(n) / n
    want = np.array([np.sum(P * np.cos(kx * r)[None, None, :]) / np.sum(P) for r in lags])
    np.testing.assert_allclose(M.sample_correlation_x(f, lags), want, atol=0.03)
    assert want[3] > np.exp(-1.0)  # the cut tail: above the continuum's exp(-r/a)


def test_vdat_has_matlab_column_order(tmp_path):
    import sgfd

    nx, ny, nz = 5, 4, 3
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    V = (100 * i + 10 * j + k).astype(np.float32)  # V(i,j,k), MATLAB order
    p = os.path.join(tmp_path, "V.dat")
    V.ravel(order="F").astype("<f4").tofile(p)  # fwrite(Fid, V, 'float')
    arr = sgfd.read_vdat(p, nx, ny, nz)
    assert arr.shape == (nz, ny, nx)
    assert np.array_equal(arr, np.transpose(V, (2, 1, 0)))
    q = os.path.join(tmp_path, "W.dat")
    sgfd.write_vdat(q, arr)
    assert open(q, "rb").read() == open(p, "rb").read()
    with pytest.raises(sgfd.SgError):
        sgfd.read_vdat(p, nx, ny, nz + 1)


def test_random_medium_rejects_bad_arguments():
    import sgfd

    for kw in ({"kind": "von_karman", "order": 0.0}, {"a": (1.0, 0.0, 1.0)}, {"eps": 0.0}, {"l": 0.0}):
        args = dict(kind="reference", a=(2.0, 2.0, 2.0), order=1.0, l=3.0, seed=0, v0=0.0, eps=0.1)
        args.update(kw)
        with pytest.raises(sgfd.SgError):
            sgfd.random_medium(8, 8, 8, **args)


def test_spectrum_kind_by_name_or_id():
    for name, kid in M.KINDS.items():
        assert np.array_equal(M.spectrum(5, 4, 3, name, (2.0, 3.0, 1.5), 0.7, 9.0),
                              M.spectrum(5, 4, 3, kid, (2.0, 3.0, 1.5), 0.7, 9.0))
    with pytest.raises(ValueError):
        M.spectrum(5, 4, 3, "gauss")
