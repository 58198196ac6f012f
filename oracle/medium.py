"""numpy restatement of the reference's random-medium builder.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ and the CPU-baseline leg of
``bench.py`` as the checker of ``sg_random_medium``; the product path never
imports this module.

Follows ``randomMedium/RandomMedia_3_d.m``:

* spectrum ``fm = (a b c) / (1 + (kx^2 a^2)^p + (ky^2 b^2)^p + (kz^2 c^2)^p)``
  with ``k = index / l`` (:20-31; the script squares ``b`` for z, with
  a = b = c = 10, p = 10, l = 50);
* random phases ``Psi = sqrt(fm) exp(i Theta)``, Theta uniform on [0, 2 pi)
  (:47-57);
* ``RealRandom = abs(ifftn(Psi))`` (:65-66);
* ``Delta = eps (RealRandom - mean) / std`` with the population standard
  deviation ``sqrt(sum(Delta0^2) / N)`` and eps = 0.1 (:74-85);
* ``V = v0 (Delta + 1)``, v0 = 2000 (:87-93).

The script draws Theta from MATLAB's unseeded ``rand``; here (as in the GPU
builder) Theta is ``2 pi u`` with ``u`` the top 53 bits of splitmix64 of
``splitmix64(seed) ^ bin`` over 2^53.  The script's spectrum is centred on
bin L/2 rather than bin 0; that shift multiplies the field by a unit-modulus
phase ramp which ``abs`` removes, so indexing by signed frequency (bin m ->
m, or m - n above (n-1)/2) gives the same |field|.  The Gaussian and von
Karman spectra (standard forms, k in radians per cell) take Re of the field.

Pins (tests/test_oracle_medium.py): splitmix64 against published outputs,
the spectrum's closed forms at chosen bins, the inverse transform against a
brute-force inverse DFT on a small grid, the normalisation invariants, and
the sample correlation of the Gaussian (exp(-r^2/a^2)) and order-1/2 von
Karman (exp(-r/a)) media.
"""
from __future__ import annotations

import numpy as np

KIND_REFERENCE, KIND_GAUSSIAN, KIND_VON_KARMAN = 0, 1, 2
KINDS = {"reference": KIND_REFERENCE, "gaussian": KIND_GAUSSIAN, "von_karman": KIND_VON_KARMAN}

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    """One splitmix64 output for state x (x += golden gamma, then the mix)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def phases(n, seed):
    """Theta of bins 0..n-1 (linear, x fastest)."""
    s0 = splitmix64(np.uint64(seed))
    u = (splitmix64(s0 ^ np.arange(n, dtype=np.uint64)) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    return 2.0 * np.pi * u


def signed_index(n):
    m = np.arange(n)
    return np.where(m <= (n - 1) // 2, m, m - n).astype(np.float64)


def kind_id(kind):
    """Spectrum kind by name ("reference", "gaussian", "von_karman") or id."""
    k = KINDS.get(kind) if isinstance(kind, str) else kind
    if k not in KINDS.values():
        raise ValueError(f"unknown spectrum kind {kind!r}")
    return k


def spectrum(nx, ny, nz, kind, a=(10.0, 10.0, 10.0), order=10.0, l=50.0):
    """P on the (nz, ny, nx) bin grid; kind by name or id."""
    kind = kind_id(kind)
    sx, sy, sz = signed_index(nx), signed_index(ny), signed_index(nz)
    if kind == KIND_REFERENCE:
        kx, ky, kz = sx / l, sy / l, sz / l
        tx = (kx * kx * a[0] * a[0]) ** order
        ty = (ky * ky * a[1] * a[1]) ** order
        tz = (kz * kz * a[2] * a[2]) ** order
        return (a[0] * a[1] * a[2]) / (1.0 + tx[None, None, :] + ty[None, :, None] + tz[:, None, None])
    kx, ky, kz = 2 * np.pi * sx / nx, 2 * np.pi * sy / ny, 2 * np.pi * sz / nz
    q = (kx * kx * a[0] * a[0])[None, None, :] + (ky * ky * a[1] * a[1])[None, :, None] + \
        (kz * kz * a[2] * a[2])[:, None, None]
    if kind == KIND_GAUSSIAN:
        return np.exp(-0.25 * q)
    return (1.0 + q) ** (-(order + 1.5))


def psi_grid(nx, ny, nz, kind, a=(10.0, 10.0, 10.0), order=10.0, l=50.0, seed=0):
    """sqrt(P) e^{i theta} on the bin grid, rounded to complex64 as the GPU
    stores it."""
    P = spectrum(nx, ny, nz, kind, a, order, l)
    th = phases(nx * ny * nz, seed).reshape(nz, ny, nx)
    amp = np.sqrt(P)
    return ((amp * np.cos(th)).astype(np.float32) + 1j * (amp * np.sin(th)).astype(np.float32)).astype(np.complex128)


def raw_field(nx, ny, nz, kind, a=(10.0, 10.0, 10.0), order=10.0, l=50.0, seed=0):
    """The field before normalisation: |ifftn(psi)| (reference) or Re (others),
    unnormalised like hipFFT's backward transform; FP64."""
    f = np.fft.ifftn(psi_grid(nx, ny, nz, kind, a, order, l, seed)) * (nx * ny * nz)
    return np.abs(f) if kind_id(kind) == KIND_REFERENCE else f.real


def normalise(x, eps=0.1, v0=2000.0):
    d = x - x.mean()
    d = eps * d / np.sqrt(np.sum(d * d) / d.size)
    return v0 * (1.0 + d) if v0 > 0 else d


def random_medium(nx, ny, nz, kind="reference", a=(10.0, 10.0, 10.0), order=10.0, l=50.0, seed=0, v0=2000.0,
                  eps=0.1):
    """FP64 result, shape (nz, ny, nx)."""
    return normalise(raw_field(nx, ny, nz, kind, a, order, l, seed), eps, v0)


def brute_inverse_dft(psi):
    """sum_k psi[k] exp(+2 pi i k.x / n), by explicit summation (small grids)."""
    nz, ny, nx = psi.shape
    out = np.zeros(psi.shape, np.complex128)
    kz, ky, kx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                ph = 2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz)
                out[z, y, x] = np.sum(psi * np.exp(ph))
    return out


def sample_correlation_x(f, lags):
    """Mean over the grid of f(x) f(x + r) / var along x (periodic)."""
    g = f - f.mean()
    v = np.mean(g * g)
    return np.array([np.mean(g * np.roll(g, -r, axis=2)) / v for r in lags])
