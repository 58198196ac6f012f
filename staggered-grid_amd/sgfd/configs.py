"""Synthetic workloads: grids, media, sources and receivers of the five stated
solver configurations, built with the library's host numerics.

Each builder returns a ``Workload``; ``solver()`` makes a configured
``sgfd.Solver`` from it.  Grid sizes are TOTAL extents (PML and the 5-cell
stencil halo included); absolute indices are used for sources and receivers.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import Solver, damping_profile, diff_coef, scales_for_maxabs, gaussian_field, ricker

HALO = 5


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    pml: int
    dt: float
    dx: float
    dy: float
    dz: float
    nt: int
    b: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    src: tuple
    wavelet: np.ndarray
    stations: np.ndarray
    comps: tuple = ("vx", "vy", "vz")
    interval: int = 1
    damp_v: float = 3000.0
    coef: np.ndarray = field(default_factory=lambda: diff_coef("exact"))
    vp_max: float = 0.0
    description: str = ""
    f0: float = 0.0
    _scales: object = None

    @property
    def points(self) -> int:
        return self.nx * self.ny * self.nz

    def damping(self):
        return [damping_profile(n, self.pml, h, "quadratic", self.damp_v)
                for n, h in ((self.nx, self.dx), (self.ny, self.dy), (self.nz, self.dz))]

    def courant(self) -> float:
        """dt*Vmax*sqrt(sum 1/h^2)*sum|c|; the leapfrog scheme is stable below 1."""
        s = math.sqrt(1 / self.dx ** 2 + 1 / self.dy ** 2 + 1 / self.dz ** 2)
        return self.dt * self.vp_max * s * float(np.sum(np.abs(self.coef)))

    def pilot_scales(self, steps=None, checks=8):
        """FP16/BF16 storage scales from an FP32 pilot run: the largest
        |value| of each field over the first `steps` steps (default 2.5
        source periods, which covers the source peak), mapped below 2^11."""
        if steps is None:
            f0 = self.f0 if self.f0 > 0 else 15.0
            steps = int(np.ceil(2.5 / (f0 * self.dt)))
        steps = max(1, min(steps, self.nt))
        s = self.solver("fp32", time_length=steps)
        try:
            m = np.zeros(9, np.float32)
            done = 0
            for c in range(checks):
                n = (steps * (c + 1)) // checks - done
                if n > 0:
                    s.step(n)
                    done += n
                m = np.maximum(m, s.field_maxabs())
        finally:
            s.close()
        return scales_for_maxabs(m)

    def solver(self, precision="fp32", kernel=None, source_mode="transient", time_length=None, scales=None):
        """A solver for this workload.  FP16/BF16 storage scales come from
        `scales` or, if None, from an FP32 pilot run (cached)."""
        if precision in ("fp16", "bf16") and scales is None:
            if self._scales is None:
                self._scales = self.pilot_scales()
            scales = self._scales
        s = Solver(self.nx, self.ny, self.nz, self.pml, self.dt, self.dx, self.dy, self.dz, coef=self.coef,
                   time_length=self.nt if time_length is None else time_length, precision=precision,
                   kernel=kernel, source_mode=source_mode)
        try:
            s.set_medium(self.b, self.lam, self.mu)
            s.set_damping(*self.damping())
            s.set_source(*self.src, self.wavelet)
            if len(self.stations):
                s.set_stations(self.stations, comps=self.comps, interval=self.interval)
            if scales is not None and precision in ("fp16", "bf16"):
                s.set_field_scales(scales)
        except Exception:
            s.close()
            raise
        return s


def elastic_from_velocity(vp, vs, rho):
    """b = 1/rho, mu = rho vs^2, lambda = rho vp^2 - 2 mu (float32)."""
    vp = np.asarray(vp, np.float64)
    vs = np.asarray(vs, np.float64)
    rho = np.asarray(rho, np.float64)
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    return (1 / rho).astype(np.float32), lam.astype(np.float32), mu.astype(np.float32)


def _surface_grid(nx, ny, z, pml, n):
    lo = HALO + pml
    xs = np.round(np.linspace(lo, nx - HALO - pml - 1, n)).astype(np.int32)
    ys = np.round(np.linspace(lo, ny - HALO - pml - 1, n)).astype(np.int32)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, z, np.int32)], axis=1).astype(np.int32)


def layered_vp(nz, ny, nx, interfaces, velocities):
    z = np.arange(nz)
    layer = np.searchsorted(np.asarray(interfaces), z, side="right")
    v = np.asarray(velocities, np.float64)[layer]
    return np.broadcast_to(v[:, None, None], (nz, ny, nx))


def plain100() -> Workload:
    """100^3 incl. a 10-cell PML, dx 10 m, dt 1 ms, 500 steps, homogeneous
    Vp 3000 / Vs 1732 / rho 2000, Ricker 15 Hz delayed 1/f0 at the centre on
    sxx/syy/szz, 20 stations on a line at z=20 recording vx/vy/vz."""
    n, pml, nt = 100, 10, 500
    shape = (n, n, n)
    b, lam, mu = elastic_from_velocity(np.full(shape, 3000.0), np.full(shape, 1732.0), np.full(shape, 2000.0))
    st = np.array([[20 + 3 * r, 50, 20] for r in range(20)], np.int32)
    return Workload("plain100", n, n, n, pml, 1e-3, 10.0, 10.0, 10.0, nt, b, lam, mu, (50, 50, 50),
                    ricker(15.0, 1e-3, nt, "none"), st, damp_v=3000.0, vp_max=3000.0,
                    description="100^3 P10 homogeneous, 500 steps", f0=15.0)


def random256(seed=1234, nt=1000) -> Workload:
    """256^3 incl. a 16-cell PML, dx 10 m, dt 1 ms, 1000 steps; Vp = 3000 (1 + 0.05 g),
    g seeded Gaussian-correlated (a = 8 cells, clipped at 4 sigma), Vs = Vp/sqrt 3,
    rho = 310 Vp^0.25; Ricker 15 Hz at (128,128,40); 64x64 receivers on z=21."""
    n, pml = 256, 16
    g = gaussian_field(n, n, n, 8.0, seed, 4.0)
    vp = 3000.0 * (1 + 0.05 * g.astype(np.float64))
    b, lam, mu = elastic_from_velocity(vp, vp / math.sqrt(3.0), 310.0 * vp ** 0.25)
    st = _surface_grid(n, n, HALO + pml, pml, 64)
    return Workload("random256", n, n, n, pml, 1e-3, 10.0, 10.0, 10.0, nt, b, lam, mu, (128, 128, 40),
                    ricker(15.0, 1e-3, nt, "none"), st, damp_v=3000.0, vp_max=float(vp.max()),
                    description="256^3 P16 random medium, 1000 steps", f0=15.0)


def layered512(seed=2024, nt=2000, nz=384) -> Workload:
    """512x512x384 incl. a 20-cell PML, dx 10 m, dt 0.8 ms, 2000 steps; four layers
    Vp 2000/3000/3800/4500 with interfaces at z=80/180/280, each with a seeded 5 %
    Gaussian-correlated perturbation (a = 8 cells, 4 sigma clip); Vs = Vp/1.75,
    Gardner density; Ricker 12 Hz at (256,256,40); 64x64 receivers on z=25."""
    nx = ny = 512
    pml = 20
    g = gaussian_field(nx, ny, nz, 8.0, seed, 4.0)
    vp = layered_vp(nz, ny, nx, (80, 180, 280), (2000.0, 3000.0, 3800.0, 4500.0)) * (1 + 0.05 * g.astype(np.float64))
    b, lam, mu = elastic_from_velocity(vp, vp / 1.75, 310.0 * vp ** 0.25)
    vmax = float(vp.max())
    del g, vp
    st = _surface_grid(nx, ny, HALO + pml, pml, 64)
    return Workload("layered512", nx, ny, nz, pml, 8e-4, 10.0, 10.0, 10.0, nt, b, lam, mu, (256, 256, 40),
                    ricker(12.0, 8e-4, nt, "none"), st, damp_v=vmax, vp_max=vmax,
                    description="512x512x384 P20 4-layer + 5% random, 2000 steps", f0=12.0)


def large1024(seed=4096, nt=1500) -> Workload:
    """1024x1024x640 incl. a 24-cell PML, dx 12.5 m, dt 1 ms, 1500 steps; layered +
    random as layered512 with interfaces at z=133/300/467; Ricker 12 Hz; 256x256
    surface receivers every 4 steps."""
    nx = ny = 1024
    nz = 640
    pml = 24
    g = gaussian_field(nx, ny, nz, 8.0, seed, 4.0)
    vp = layered_vp(nz, ny, nx, (133, 300, 467), (2000.0, 3000.0, 3800.0, 4500.0)) * (1 + 0.05 * g.astype(np.float64))
    del g
    b, lam, mu = elastic_from_velocity(vp, vp / 1.75, 310.0 * vp ** 0.25)
    vmax = float(vp.max())
    del vp
    st = _surface_grid(nx, ny, HALO + pml, pml, 256)
    return Workload("large1024", nx, ny, nz, pml, 1e-3, 12.5, 12.5, 12.5, nt, b, lam, mu, (512, 512, 60),
                    ricker(12.0, 1e-3, nt, "none"), st, interval=4, damp_v=vmax, vp_max=vmax,
                    description="1024x1024x640 P24 layered + random, 1500 steps", f0=12.0)


def batched_shots(seed=77, nshots=32, nt=800):
    """32 independent shots: nx, ny ~ U{128..320}, nz ~ U{96..256}, 12-cell PML,
    dx 10 m, dt 1 ms, 800 steps; per shot a seeded random medium (mean Vp
    U[2500,3500], 5 %, a = 6 cells), Ricker f0 ~ U[10,20] Hz at a random interior
    point >= 10 cells from the PML."""
    rng = np.random.default_rng(seed)
    shots = []
    for s in range(nshots):
        nx, ny = (int(v) for v in rng.integers(128, 321, 2))
        nz = int(rng.integers(96, 257))
        pml = 12
        vmean = float(rng.uniform(2500.0, 3500.0))
        f0 = float(rng.uniform(10.0, 20.0))
        lo = HALO + pml + 10
        src = tuple(int(rng.integers(lo, n - lo)) for n in (nx, ny, nz))
        g = gaussian_field(nx, ny, nz, 6.0, seed * 1000 + s, 4.0)
        vp = vmean * (1 + 0.05 * g.astype(np.float64))
        b, lam, mu = elastic_from_velocity(vp, vp / math.sqrt(3.0), 310.0 * vp ** 0.25)
        st = _surface_grid(nx, ny, HALO + pml, pml, 16)
        shots.append(Workload(f"shot{s}", nx, ny, nz, pml, 1e-3, 10.0, 10.0, 10.0, nt, b, lam, mu, src,
                              ricker(f0, 1e-3, nt, "none"), st, damp_v=float(vp.max()), vp_max=float(vp.max()),
                              description=f"shot {s}: {nx}x{ny}x{nz}", f0=f0))
    return shots


def nopml500(nt=1000) -> Workload:
    """The configuration of the reference's published GPU timing
    (3DGPUwithoutPML_1DArraySingleGPU/main.cu:11-41): 500^3 with no PML
    (5-cell halo only), dx 5 m, dt 0.5 ms, 1000 steps, homogeneous rho 2500,
    lambda = rho 2000^2, mu = lambda/3, Ricker 30 Hz at the centre."""
    n = 500
    shape = (n, n, n)
    rho = np.float32(2500.0)
    lam = np.float32(2500.0 * 2000.0 ** 2)
    b = np.full(shape, np.float32(1.0) / rho, np.float32)
    lam_a = np.full(shape, lam, np.float32)
    mu_a = np.full(shape, np.float32(lam / np.float32(3.0)), np.float32)
    vp = float(np.sqrt((lam + 2 * lam / 3) / rho))
    st = np.array([[n // 2 + 20, n // 2, n // 2], [n // 2, n // 2 + 20, n // 2]], np.int32)
    return Workload("nopml500", n, n, n, 0, 5e-4, 5.0, 5.0, 5.0, nt, b, lam_a, mu_a, (n // 2, n // 2, n // 2),
                    ricker(30.0, 5e-4, nt, "none"), st, damp_v=vp, vp_max=vp,
                    description="500^3 no PML, 1000 steps (the reference's published GPU timing case)", f0=30.0)


BUILDERS = {"plain100": plain100, "random256": random256, "layered512": layered512, "large1024": large1024,
            "nopml500": nopml500}
