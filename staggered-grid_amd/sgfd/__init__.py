"""Python host layer over the sgfd C ABI (libsgfd.so, include/sgfd.h).

``Solver`` is a thin object wrapper of the C entry points; ``ElasticWaveEquation``
mirrors the reference driver class (3DGPUwithPML_1DArraySingleGPU_concise/
ElasticWaveEquation.h:17-60) so a reference user finds the same constructor,
``run()``, ``propagate()``, ``sourceFunction()`` and ``stationDataOutput()``.

All compute runs in the HIP library; if ``libsgfd.so`` is missing this module
raises at import of the library (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os
from collections import namedtuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SGFD_LIB selects another build of the library (used for build-variant experiments)
LIB_PATH = os.environ.get("SGFD_LIB", os.path.join(_HERE, "libsgfd.so"))

FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
PRECISIONS = {"fp32": 0, "fp64": 1, "fp16": 2, "bf16": 3}
SOURCE_MODES = {"transient": 0, "accumulate": 1}
KERNELS = {"tiled": 0, "simple": 1}
DAMP_REFERENCE, DAMP_QUADRATIC, DAMP_CUDA = 0, 1, 2
RICKER_CEIL_DELAY, RICKER_NO_DELAY = 0, 1


class SgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sgfd error {code}: {msg}")
        self.code = code


class _Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("pml", ctypes.c_int),
        ("dt", ctypes.c_float), ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
        ("diff_coef", ctypes.c_float * 5),
        ("time_length", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("source_mode", ctypes.c_int),
        ("kernel", ctypes.c_int),
    ]


_lib = None
_FP = ctypes.POINTER(ctypes.c_float)
class _Tile(ctypes.Structure):
    _fields_ = [("tile_y", ctypes.c_int), ("rows_per_thread", ctypes.c_int), ("pml_tile_y", ctypes.c_int),
                ("pml_rows_per_thread", ctypes.c_int), ("z_chunk", ctypes.c_int), ("waves_per_simd", ctypes.c_int)]

    def as_tuple(self):
        return (self.tile_y, self.rows_per_thread, self.pml_tile_y, self.pml_rows_per_thread, self.z_chunk,
                self.waves_per_simd)


_TP = ctypes.POINTER(_Tile)


class _Spectrum(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("a", ctypes.c_double * 3), ("order", ctypes.c_double),
                ("l", ctypes.c_double)]


class _Stability(ctypes.Structure):
    _fields_ = [("courant_limit", ctypes.c_double), ("check_interval", ctypes.c_int), ("max_growth", ctypes.c_double)]


SPECTRA = {"reference": 0, "gaussian": 1, "von_karman": 2}
ERR_UNSTABLE = -7
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_LLP = ctypes.POINTER(ctypes.c_longlong)
_VP = ctypes.c_void_p

# exported symbol -> (restype, argtypes); the list is the C ABI of include/sgfd.h
API = {
    "sg_api_version": (ctypes.c_int, []),
    "sg_last_error": (ctypes.c_char_p, []),
    "sg_create": (ctypes.c_int, [ctypes.POINTER(_Config), ctypes.POINTER(_VP)]),
    "sg_destroy": (ctypes.c_int, [_VP]),
    "sg_set_medium": (ctypes.c_int, [_VP, _FP, _FP, _FP]),
    "sg_set_medium_velocity": (ctypes.c_int, [_VP, _FP, _FP, _FP]),
    "sg_set_damping": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "sg_set_source": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _FP, ctypes.c_int]),
    "sg_set_source_sample": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_float]),
    "sg_set_stations": (ctypes.c_int, [_VP, _IP, ctypes.c_int, ctypes.c_uint, ctypes.c_int]),
    "sg_set_field_scales": (ctypes.c_int, [_VP, _FP]),
    "sg_get_field_scales": (ctypes.c_int, [_VP, _FP]),
    "sg_step": (ctypes.c_int, [_VP, ctypes.c_int]),
    "sg_run": (ctypes.c_int, [_VP]),
    "sg_synchronize": (ctypes.c_int, [_VP]),
    "sg_reset": (ctypes.c_int, [_VP]),
    "sg_steps_done": (ctypes.c_longlong, [_VP]),
    "sg_get_stations": (ctypes.c_int, [_VP, _FP, ctypes.c_longlong, _LLP]),
    "sg_get_field": (ctypes.c_int, [_VP, ctypes.c_int, _FP]),
    "sg_get_field_f64": (ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    "sg_energy": (ctypes.c_int, [_VP, _DP]),
    "sg_energy_full": (ctypes.c_int, [_VP, _DP]),
    "sg_energy_noalloc": (ctypes.c_int, [_VP, _DP]),
    "sg_courant": (ctypes.c_int, [_VP, _DP]),
    "sg_courant_number": (ctypes.c_int, [ctypes.c_longlong, _FP, _FP, _FP, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_float, _FP, _DP]),
    "sg_set_stability": (ctypes.c_int, [_VP, ctypes.POINTER(_Stability)]),
    "sg_get_stability": (ctypes.c_int, [_VP, ctypes.POINTER(_Stability)]),
    "sg_field_maxabs": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "sg_set_rounding": (ctypes.c_int, [_VP, ctypes.c_int]),
    "sg_set_damping_local": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "sg_set_streams": (ctypes.c_int, [_VP, ctypes.c_int]),
    "sg_get_streams": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
    "sg_set_compensation": (ctypes.c_int, [_VP, ctypes.c_uint]),
    "sg_step_timed": (ctypes.c_int, [_VP, ctypes.c_int, _DP, _DP, _DP]),
    "sg_set_graph": (ctypes.c_int, [_VP, ctypes.c_int]),
    "sg_set_tile": (ctypes.c_int, [_VP, _TP]),
    "sg_get_tile": (ctypes.c_int, [_VP, _TP]),
    "sg_sweep_bytes": (ctypes.c_int, [_VP, _DP, _DP]),
    "sg_autotune": (ctypes.c_int, [_VP, ctypes.c_int, _TP, _DP]),
    "sg_copy_bandwidth": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_int, _DP]),
    "sg_batch_create": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.POINTER(_VP)]),
    "sg_batch_set_launch": (ctypes.c_int, [_VP, _TP, ctypes.c_int, ctypes.c_int]),
    "sg_batch_step": (ctypes.c_int, [_VP, ctypes.c_int]),
    "sg_batch_step_timed": (ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    "sg_batch_synchronize": (ctypes.c_int, [_VP]),
    "sg_batch_autotune": (ctypes.c_int, [_VP, ctypes.c_int, _TP, _IP, _DP]),
    "sg_batch_destroy": (ctypes.c_int, [_VP]),
    "sg_ricker": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, _FP]),
    "sg_damping_profile": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, _DP]),
    "sg_diff_coef": (ctypes.c_int, [ctypes.c_int, _FP]),
    "sg_gaussian_field": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_ulonglong, ctypes.c_double, _FP]),
    "sg_read_raw_f32": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_longlong, _FP]),
    "sg_elastic_from_velocity": (ctypes.c_int, [ctypes.c_longlong, _FP, _FP, _FP, _FP, _FP, _FP]),
    "sg_write_raw_f32": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_longlong, _FP]),
    "sg_random_medium": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Spectrum),
                                        ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double, _FP]),
    "sg_random_medium_timed": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Spectrum),
                                              ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double, _FP, _DP]),
}


def lib():
    """Load libsgfd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsgfd.so not built at {LIB_PATH}: run __graft_entry__.build() "
                               "(make -C staggered-grid_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in API.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SgError(rc, lib().sg_last_error().decode())


def _fptr(a):
    return a.ctypes.data_as(_FP)


def diff_coef(kind: str = "reference") -> np.ndarray:
    """Staggered coefficients: 'reference' literals or 'exact' rationals."""
    out = np.zeros(5, np.float32)
    _check(lib().sg_diff_coef({"reference": 0, "exact": 1}[kind], _fptr(out)))
    return out


def ricker(f0: float, dt: float, nt: int, delay: str = "ceil") -> np.ndarray:
    out = np.zeros(nt, np.float32)
    _check(lib().sg_ricker(f0, dt, nt, RICKER_CEIL_DELAY if delay == "ceil" else RICKER_NO_DELAY, _fptr(out)))
    return out


def damping_profile(n, pml, h, kind="reference", v=2000.0, r=1e-3, tau=3.4) -> np.ndarray:
    out = np.zeros(n, np.float64)
    k = {"reference": DAMP_REFERENCE, "quadratic": DAMP_QUADRATIC, "cuda": DAMP_CUDA}[kind]
    _check(lib().sg_damping_profile(n, pml, h, k, v, r, tau, out.ctypes.data_as(_DP)))
    return out


def gaussian_field(nx, ny, nz, corr, seed, clip=4.0) -> np.ndarray:
    """Seeded unit-variance field with correlation exp(-r^2/corr^2), shape (nz, ny, nx)."""
    out = np.zeros((nz, ny, nx), np.float32)
    _check(lib().sg_gaussian_field(nx, ny, nz, float(corr), int(seed), float(clip), _fptr(out)))
    return out


def elastic_from_velocity(vp, vs, rho):
    """(b, lambda, mu) as float32 arrays of the input shape: b = 1/rho,
    mu = rho vs^2, lambda = rho vp^2 - 2 mu, formed in double by the library."""
    arrs = [np.ascontiguousarray(a, np.float32) for a in (vp, vs, rho)]
    if not (arrs[0].shape == arrs[1].shape == arrs[2].shape):
        raise ValueError("vp, vs and rho must have one shape")
    out = [np.empty_like(arrs[0]) for _ in range(3)]
    _check(lib().sg_elastic_from_velocity(arrs[0].size, *[_fptr(a) for a in arrs], *[_fptr(o) for o in out]))
    return tuple(out)


def courant_number(b, lam, mu, dt, dx, dy, dz, coef) -> float:
    """Normalised Courant number of a model (stable at <= 1): dt * max(vp) *
    sum|c_n| * sqrt(1/dx^2 + 1/dy^2 + 1/dz^2), computed by the library on the
    host."""
    arrs = [np.ascontiguousarray(a, np.float32).ravel() for a in (b, lam, mu)]
    c = np.ascontiguousarray(coef, np.float32)
    out = ctypes.c_double()
    _check(lib().sg_courant_number(arrs[0].size, *[_fptr(a) for a in arrs], dt, dx, dy, dz, _fptr(c),
                                   ctypes.byref(out)))
    return out.value


def read_vdat(path, nx, ny, nz) -> np.ndarray:
    """Raw float32 model file as written by the reference's MATLAB tools
    (randomMedium/RandomMedia_3_d.m:101-103: fwrite of V(i,j,k), i fastest),
    returned as an (nz, ny, nx) array with x = i."""
    out = np.empty((nz, ny, nx), np.float32)
    _check(lib().sg_read_raw_f32(os.fsencode(path), out.size, _fptr(out)))
    return out


def write_vdat(path, v) -> None:
    """Write an (nz, ny, nx) array as raw float32, x fastest (MATLAB's
    fread(fid, [nx*ny*nz], 'float') order for V(i,j,k))."""
    a = np.ascontiguousarray(v, np.float32)
    _check(lib().sg_write_raw_f32(os.fsencode(path), a.size, _fptr(a)))


def random_medium(nx, ny, nz, kind="reference", a=(10.0, 10.0, 10.0), order=10.0, l=50.0, seed=0,
                  v0=2000.0, eps=0.1) -> np.ndarray:
    """Seeded random medium by the reference's spectral method on the GPU
    (randomMedium/RandomMedia_3_d.m:20-93): v0 * (1 + delta) with delta of
    zero mean and standard deviation eps (delta itself when v0 <= 0), shape
    (nz, ny, nx).  kind: "reference" (the script's spectrum; its defaults are
    a = 10, order = 10, l = 50, v0 = 2000, eps = 0.1), "gaussian" or
    "von_karman" (a in cells; order = von Karman order)."""
    out, _ = random_medium_timed(nx, ny, nz, kind, a, order, l, seed, v0, eps)
    return out


def random_medium_timed(nx, ny, nz, kind="reference", a=(10.0, 10.0, 10.0), order=10.0, l=50.0, seed=0,
                        v0=2000.0, eps=0.1):
    """random_medium plus the device time of the build in ms (FFT plan and
    the copy to the host excluded)."""
    sp = _Spectrum(SPECTRA[kind], (ctypes.c_double * 3)(*[float(x) for x in a]), float(order), float(l))
    out = np.empty((nz, ny, nx), np.float32)
    ms = ctypes.c_double()
    _check(lib().sg_random_medium_timed(nx, ny, nz, ctypes.byref(sp), int(seed), float(v0), float(eps), _fptr(out),
                                        ctypes.byref(ms)))
    return out, ms.value


def copy_bandwidth(nbytes: int, reps: int = 20) -> float:
    g = ctypes.c_double()
    _check(lib().sg_copy_bandwidth(int(nbytes), reps, ctypes.byref(g)))
    return g.value


class Solver:
    """One shot on the current GPU.  Arrays are numpy, x-fastest, shape (nz, ny, nx)."""

    def __init__(self, nx, ny, nz, pml, dt, dx, dy, dz, coef=None, time_length=0, precision="fp32",
                 source_mode="transient", kernel=None):
        # kernel None: the tiled sweeps, or for FP64 (the checking build,
        # which the library runs on the per-point kernels only) "simple"
        if kernel is None:
            kernel = "simple" if precision == "fp64" else "tiled"
        c = _Config()
        c.nx, c.ny, c.nz, c.pml = nx, ny, nz, pml
        c.dt, c.dx, c.dy, c.dz = dt, dx, dy, dz
        coef = diff_coef("reference") if coef is None else np.asarray(coef, np.float32)
        for i in range(5):
            c.diff_coef[i] = float(coef[i])
        c.time_length = time_length
        c.precision = PRECISIONS[precision]
        c.source_mode = SOURCE_MODES[source_mode]
        c.kernel = KERNELS[kernel]
        h = _VP()
        _check(lib().sg_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.shape = (nz, ny, nx)
        self.nx, self.ny, self.nz, self.pml = nx, ny, nz, pml
        self.dt = dt
        self.time_length = time_length
        self.precision = precision
        self._nst = 0
        self._ncomp = 0
        self._interval = 1

    # -- lifetime
    def close(self):
        """Frees the solver; raises SgError while a live Batch holds it."""
        if getattr(self, "_h", None):
            _check(lib().sg_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- inputs
    def set_medium(self, b, lam, mu):
        arrs = [np.ascontiguousarray(a, np.float32).reshape(self.shape) for a in (b, lam, mu)]
        _check(lib().sg_set_medium(self._h, *[_fptr(a) for a in arrs]))

    def set_medium_velocity(self, vp, vs, rho):
        """The model as P velocity, S velocity and density (x-fastest)."""
        arrs = [np.ascontiguousarray(a, np.float32).reshape(self.shape) for a in (vp, vs, rho)]
        _check(lib().sg_set_medium_velocity(self._h, *[_fptr(a) for a in arrs]))

    def set_damping(self, dx, dy, dz):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (dx, dy, dz)]
        assert arrs[0].size == self.nx and arrs[1].size == self.ny and arrs[2].size == self.nz
        _check(lib().sg_set_damping(self._h, *[a.ctypes.data_as(_DP) for a in arrs]))

    def set_damping_local(self, gx, gy, gz):
        """Per-cell damping d = g(axis index) * sqrt(mu*b) of the cell (the
        CUDA reference path's form); g from damping_profile(..., v=1)."""
        arrs = [np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz)]
        assert arrs[0].size == self.nx and arrs[1].size == self.ny and arrs[2].size == self.nz
        _check(lib().sg_set_damping_local(self._h, *[a.ctypes.data_as(_DP) for a in arrs]))

    def set_source(self, x, y, z, wavelet):
        w = np.ascontiguousarray(wavelet, np.float32)
        _check(lib().sg_set_source(self._h, int(x), int(y), int(z), _fptr(w), w.size))

    def set_source_sample(self, it, value):
        _check(lib().sg_set_source_sample(self._h, int(it), float(value)))

    def set_stations(self, xyz, comps=("vx", "vy", "vz"), interval=1):
        st = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        mask = 0
        for c in comps:
            mask |= 1 << FIELD_ID[c]
        _check(lib().sg_set_stations(self._h, st.ctypes.data_as(_IP), st.shape[0], mask, interval))
        self._nst = st.shape[0]
        self._ncomp = len([f for f in FIELDS if mask >> FIELD_ID[f] & 1])
        self._interval = interval

    def set_field_scales(self, scales):
        sc = np.ascontiguousarray(scales, np.float32)
        assert sc.size == 9
        _check(lib().sg_set_field_scales(self._h, _fptr(sc)))

    def field_scales(self):
        out = np.zeros(12, np.float32)
        _check(lib().sg_get_field_scales(self._h, _fptr(out)))
        return out

    # -- stepping
    def step(self, n=1):
        _check(lib().sg_step(self._h, int(n)))

    def run(self):
        _check(lib().sg_run(self._h))

    def synchronize(self):
        _check(lib().sg_synchronize(self._h))

    def reset(self):
        _check(lib().sg_reset(self._h))

    @property
    def steps_done(self):
        return lib().sg_steps_done(self._h)

    def step_timed(self, n):
        v, s, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().sg_step_timed(self._h, int(n), ctypes.byref(v), ctypes.byref(s), ctypes.byref(t)))
        return v.value, s.value, t.value

    def set_graph(self, enable=True):
        _check(lib().sg_set_graph(self._h, 1 if enable else 0))

    def set_tile(self, ty=4, ny=1, pty=4, pny=1, zchunk=0, waves=0):
        """Box shapes: interior (ty rows per box, ny rows per thread), PML
        (pty, pny), zchunk planes per box (0 = auto); waves: register bound
        of the one-launch-per-sweep form (0 none, 2, 3, 4)."""
        _check(lib().sg_set_tile(self._h, ctypes.byref(_Tile(ty, ny, pty, pny, zchunk, waves))))

    def tile(self):
        t = _Tile()
        _check(lib().sg_get_tile(self._h, ctypes.byref(t)))
        return t.as_tuple()

    def autotune(self, steps=5):
        """Pick the fastest box shapes, register bound and stream count
        (resets the wavefields); returns ((ty, ny, pty, pny, zchunk, waves),
        ms per step)."""
        t, ms = _Tile(), ctypes.c_double()
        _check(lib().sg_autotune(self._h, steps, ctypes.byref(t), ctypes.byref(ms)))
        return t.as_tuple(), ms.value

    def sweep_bytes(self):
        v, s = ctypes.c_double(), ctypes.c_double()
        _check(lib().sg_sweep_bytes(self._h, ctypes.byref(v), ctypes.byref(s)))
        return v.value, s.value

    # -- outputs
    def stations(self) -> np.ndarray:
        n = ctypes.c_longlong()
        _check(lib().sg_get_stations(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.float32)
        if n.value:
            _check(lib().sg_get_stations(self._h, _fptr(out), n.value, ctypes.byref(n)))
        return out.reshape(-1, self._nst, self._ncomp) if self._nst else out

    def field(self, name, dtype=np.float32) -> np.ndarray:
        fid = FIELD_ID[name] if isinstance(name, str) else int(name)
        if dtype == np.float64:
            out = np.zeros(self.shape, np.float64)
            _check(lib().sg_get_field_f64(self._h, fid, out.ctypes.data_as(_DP)))
        else:
            out = np.zeros(self.shape, np.float32)
            _check(lib().sg_get_field(self._h, fid, _fptr(out)))
        return out

    def fields(self, dtype=np.float32) -> np.ndarray:
        return np.stack([self.field(f, dtype) for f in FIELDS])

    def energy(self):
        out = np.zeros(2, np.float64)
        _check(lib().sg_energy(self._h, out.ctypes.data_as(_DP)))
        return out

    def energy_full(self):
        """(kinetic, sum of stresses squared, strain energy), FP64, summed in
        a fixed order."""
        out = np.zeros(3, np.float64)
        _check(lib().sg_energy_full(self._h, out.ctypes.data_as(_DP)))
        return out

    def courant(self) -> float:
        """Normalised Courant number of the current model (stable at <= 1)."""
        c = ctypes.c_double()
        _check(lib().sg_courant(self._h, ctypes.byref(c)))
        return c.value

    def set_stability(self, courant_limit=1.0, check_interval=0, max_growth=2.0):
        """Courant limit enforced by set_medium (0: none) and the FP64 energy
        check every check_interval steps (0: off); a failed check raises
        SgError with code ERR_UNSTABLE."""
        st = _Stability(float(courant_limit), int(check_interval), float(max_growth))
        _check(lib().sg_set_stability(self._h, ctypes.byref(st)))

    def stability(self):
        st = _Stability()
        _check(lib().sg_get_stability(self._h, ctypes.byref(st)))
        return st.courant_limit, st.check_interval, st.max_growth

    def set_streams(self, n):
        """Streams the work-lists of a sweep are spread over (1, 2 or 3)."""
        _check(lib().sg_set_streams(self._h, int(n)))

    def streams(self):
        n = ctypes.c_int()
        _check(lib().sg_get_streams(self._h, ctypes.byref(n)))
        return n.value

    def set_rounding(self, mode):
        """16-bit storage rounding: "nearest" (even) or "stochastic"."""
        _check(lib().sg_set_rounding(self._h, {"nearest": 0, "stochastic": 1}[mode]))

    def set_compensation(self, fields):
        """Compensated time update (16-bit storage) for the given fields:
        an iterable of names ("vx", ..., "syz"), "all", "velocity",
        "stress", or an int bit mask."""
        if isinstance(fields, int):
            mask = fields
        elif fields == "all":
            mask = 0x1FF
        elif fields == "velocity":
            mask = 0x7
        elif fields == "stress":
            mask = 0x1F8
        else:
            mask = 0
            for f in fields:
                mask |= 1 << FIELDS.index(f)
        _check(lib().sg_set_compensation(self._h, mask))

    def field_maxabs(self):
        """Largest |value| of each of the nine wavefields (unscaled)."""
        out = np.zeros(9, np.float32)
        _check(lib().sg_field_maxabs(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out


class Batch:
    """Several FP32 solvers advanced by one launch per half-step (sg_batch_*).
    A shot's source, stations, damping or medium may change between batch
    steps (the next step picks the change up); call synchronize() before
    stepping or reading a shot on its own."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (_VP * len(self.solvers))(*[s.handle for s in self.solvers])
        h = _VP()
        _check(lib().sg_batch_create(arr, len(self.solvers), ctypes.byref(h)))
        self._h = h

    def set_launch(self, ty=16, ny=1, pty=4, pny=1, zchunk=24, grid=0, graph=True, waves=0):
        """Box shapes, z-chunk, grid (0: one block per item), graph replay
        and the register bound; (4,1) boxes throughout run each half-step as
        one launch."""
        _check(lib().sg_batch_set_launch(self._h, ctypes.byref(_Tile(ty, ny, pty, pny, zchunk, waves)), grid,
                                         1 if graph else 0))

    def autotune(self, steps=5):
        """Pick z-chunk, register bound and grid of the batched launch
        (resets every shot); returns ((ty, ny, pty, pny, zchunk, waves),
        grid, ms per step)."""
        t, g, ms = _Tile(), ctypes.c_int(), ctypes.c_double()
        _check(lib().sg_batch_autotune(self._h, int(steps), ctypes.byref(t), ctypes.byref(g), ctypes.byref(ms)))
        return t.as_tuple(), g.value, ms.value

    def step(self, n=1):
        _check(lib().sg_batch_step(self._h, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_double()
        _check(lib().sg_batch_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        _check(lib().sg_batch_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().sg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Reference-shaped driver
# ---------------------------------------------------------------------------

GRID = namedtuple("GRID", "xLength yLength zLength")
Delta = namedtuple("Delta", "t x y z")
SOURCELOCATION = namedtuple("SOURCELOCATION", "x y z")
STATIONLOCATION = namedtuple("STATIONLOCATION", "x y z")
Medium = namedtuple("Medium", "buoyancy lambda_ mu")


class ElasticWaveEquation:
    """Mirror of ElasticWaveEquation (3DGPUwithPML_1DArraySingleGPU_concise/
    ElasticWaveEquation.h:41-58): total-size GRID (PML and 5-cell halo
    included), Delta (t, x, y, z), Medium (buoyancy, lambda, mu; x-fastest
    flat or (nz, ny, nx)), SOURCELOCATION, one STATIONLOCATION or a list of
    them, DifferenceCoefficient (5 floats) and the Ricker main frequency.

    ``gpuDim`` is accepted and ignored (launch shapes are chosen by the
    library).  Defaults follow the reference CUDA path: sxx recorded,
    accumulating source, reference damping formula; pass
    ``source_mode="transient"`` for the CPU path's behaviour and
    ``damping="cuda"`` for the CUDA path's per-cell damping (tau 4.3, local
    S velocity; ..._concise/ElasticMacro.h:21,26-27).
    """

    def __init__(self, timeLength, gpuDim, pml, grid, delta, medCPU, sourceLocation, stationLocation, dc,
                 mainFrequncy, components=("sxx",), source_mode="accumulate", precision="fp32",
                 kernel=None, damping=None):
        self.timeLength = timeLength
        self.pml = pml
        self.grid = GRID(*grid)
        self.delta = Delta(*delta)
        self.mainFrequncy = mainFrequncy
        self.sourceLocationCPU = SOURCELOCATION(*sourceLocation)
        st = np.asarray(stationLocation, np.int32).reshape(-1, 3)
        self.solver = Solver(self.grid.xLength, self.grid.yLength, self.grid.zLength, pml, self.delta.t,
                             self.delta.x, self.delta.y, self.delta.z, coef=dc, time_length=timeLength,
                             precision=precision, source_mode=source_mode, kernel=kernel)
        self.solver.set_medium(*medCPU)
        if isinstance(damping, str) and damping == "cuda":
            self.solver.set_damping_local(*[damping_profile(n, pml, h, "cuda", v=1.0, tau=4.3)
                                           for n, h in ((self.grid.xLength, self.delta.x),
                                                        (self.grid.yLength, self.delta.y),
                                                        (self.grid.zLength, self.delta.z))])
        elif damping is not None:
            self.solver.set_damping(*damping)
        self._wavelet = np.array([self.sourceFunction(it) for it in range(timeLength)], np.float32)
        self.solver.set_source(*self.sourceLocationCPU, self._wavelet)
        self.solver.set_stations(st, comps=components)
        self.stationDataCPU = None

    def sourceFunction(self, timeIndex):
        """Ricker of the reference (ElasticWaveEquation.cu:187-192), float arithmetic."""
        return float(ricker(self.mainFrequncy, self.delta.t, timeIndex + 1)[timeIndex])

    def propagate(self, timeIndex, sourceValue):
        """One step with an explicit source sample (ElasticWaveEquation.cu:194-206)."""
        if timeIndex != self.solver.steps_done:
            raise ValueError("propagate() must be called with consecutive time indices")
        self.solver.set_source_sample(timeIndex, sourceValue)
        self.solver.step(1)

    def run(self, snap_every=100, snap_from=500, snap_dir="."):
        """All timeLength steps, slices through the source every `snap_every`
        steps from `snap_from` on (the reference: every 100 from 500; 0 = no
        snapshots), then the station buffer to host
        (ElasticWaveEquation.cu:140-185)."""
        if snap_every <= 0:
            self.solver.run()
        else:
            for it in range(self.solver.steps_done, self.timeLength):
                self.solver.step(1)
                if it % snap_every == 0 and it >= snap_from:
                    self.snapShootSlice(it, snap_dir)
        self.stationDataCPU = self.solver.stations()
        return self.stationDataCPU

    @staticmethod
    def _write_lines(path, values):
        with open(path, "w") as f:
            f.write("".join("%g\n" % float(v) for v in np.asarray(values, np.float32).ravel()))

    def snapShootSlice(self, it, snap_dir="."):
        """vx, sxx and sxy on the three grid planes through the source, one
        value per line, in the reference's files and order
        (ElasticWaveEquation.cu:929-1005)."""
        sx, sy, sz = self.sourceLocationCPU
        for name, fld in (("velocityX", "vx"), ("stressXX", "sxx"), ("stressXY", "sxy")):
            F = self.solver.field(fld)  # (nz, ny, nx)
            self._write_lines(os.path.join(snap_dir, f"{name}SliceXX_YoZ_{it}.txt"), F[:, :, sx])
            self._write_lines(os.path.join(snap_dir, f"{name}SliceXX_XoZ_{it}.txt"), F[:, sy, :])
            self._write_lines(os.path.join(snap_dir, f"{name}SliceXX_XoY_{it}.txt"), F[sz, :, :])

    def snapShootVolum(self, it, snap_dir="."):
        """sxx over the updated volume [5, n-5)^3, one value per line
        (ElasticWaveEquation.cu:1007-1037)."""
        F = self.solver.field("sxx")
        self._write_lines(os.path.join(snap_dir, f"snapshot_0_{it}.txt"), F[5:-5, 5:-5, 5:-5])

    def snapShootBinary(self, it, field, snap_dir="."):
        """One field over the whole grid as raw little-endian float32 (x
        fastest), file <field>_<it>.bin; no counterpart in the reference."""
        name = field if isinstance(field, str) else FIELDS[field]
        self.solver.field(name).astype("<f4").tofile(os.path.join(snap_dir, f"{name}_{it}.bin"))

    def stationDataOutput(self, path="stationData.txt"):
        """One value per line per recorded sample (ElasticWaveEquation.cu:1040-1050)."""
        data = self.solver.stations() if self.stationDataCPU is None else self.stationDataCPU
        with open(path, "w") as f:
            for v in np.asarray(data).reshape(data.shape[0], -1):
                f.write(" ".join(f"{x:g}" for x in v) + "\n")

    def close(self):
        self.solver.close()


class ElasticWaveEquation3D:
    """Mirror of the CPU driver class (3DCPUwithPML_3DArray_GridIncludePMLandOrder/
    ElasticWaveEquation3D.h:11-52).  GRID is the interior extent; buoyancy,
    lambda_ and mu span L + 2*pml + 10 points per axis and are indexed
    [i][j][k] like the reference's float*** (numpy shape (nx, ny, nz), k
    fastest); the source is in total-grid indices; the defaults are the
    reference's (5 m, 0.5 ms, 30 Hz).  The numerics are the CPU reference's:
    the DIFF_10 literals, its damping formula (tau 3.4, Vs 2000, R 0.001 as
    float constants), the Ricker with the ceil delay and the transient source.
    propagate() writes curveLine1..3.txt, snapshot_<q>_<it>.txt and
    dampingX/Y/Z.txt (ElasticWaveEquation3D.cpp:53-112,302-311) into out_dir,
    the same text as include/ElasticWaveEquation3D.hpp writes."""

    def __init__(self, grid, timeLength, pml, buoyancy, lambda_, mu, sourceLocation, deltaX=5.0, deltaY=5.0,
                 deltaZ=5.0, deltaT=0.0005, mainFrequncy=30.0, precision="fp32", kernel=None):
        self.grid = GRID(*grid)
        self.timeLength = int(timeLength)
        self.pml = int(pml)
        self.n = tuple(int(L) + 2 * self.pml + 10 for L in self.grid)
        self.model = (buoyancy, lambda_, mu)
        for a in self.model:
            if np.shape(a) != self.n:
                raise ValueError(f"model arrays must have shape (nx, ny, nz) = {self.n}")
        self.sourceLocation = SOURCELOCATION(*sourceLocation)
        self.deltaT = deltaT
        self.mainFrequncy = mainFrequncy
        f32 = lambda x: float(np.float32(x))
        nx, ny, nz = self.n
        self.solver = Solver(nx, ny, nz, self.pml, deltaT, deltaX, deltaY, deltaZ, coef=diff_coef("reference"),
                             time_length=self.timeLength, precision=precision, source_mode="transient",
                             kernel=kernel)
        self.damping = [damping_profile(m, self.pml, f32(h), "reference", f32(2000.0), f32(0.001), f32(3.4))
                        for m, h in zip(self.n, (deltaX, deltaY, deltaZ))]
        self.solver.set_damping(*self.damping)
        if self.timeLength > 0:
            self.solver.set_source(*self.sourceLocation, ricker(mainFrequncy, deltaT, self.timeLength, "ceil"))
        e = self.pml + 10
        sx, sy, sz = self.sourceLocation
        self.solver.set_stations(np.array([[e, sy, sz], [sx, e, sz], [sx, sy, e]], np.int32), comps=("sxx",))
        self.curveLineData = None
        self._propagated = False

    def sourceFunction(self, timeIndex):
        """Ricker of the reference (ElasticWaveEquation3D.cpp:114-119)."""
        return float(ricker(self.mainFrequncy, self.deltaT, timeIndex + 1, "ceil")[timeIndex])

    def propagate(self, out_dir=".", write_files=True):
        """The time loop (ElasticWaveEquation3D.cpp:53-85); returns the three
        curves.  Runs once per object (the reference would restart its time
        index on the live fields)."""
        if self._propagated:
            raise RuntimeError("propagate() runs once per object")
        self._propagated = True
        self.solver.set_medium(*[np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0))
                                 for a in self.model])
        if write_files:
            self.calculateDamping(out_dir)
        done = 0
        for it in range(self.timeLength):
            if ((it + 1) >= 200 and it % 50 == 0) or it + 1 == 300:
                self.solver.step(it + 1 - done)
                done = it + 1
                if write_files:
                    self.snapshoot(it + 1, out_dir)
        if done < self.timeLength:
            self.solver.step(self.timeLength - done)
        tr = self.solver.stations().reshape(-1, 3)
        self.curveLineData = [tr[:, c].copy() for c in range(3)]
        if write_files:
            for c in range(3):
                ElasticWaveEquation._write_lines(os.path.join(out_dir, f"curveLine{c + 1}.txt"), tr[:, c])
        return self.curveLineData

    def snapshoot(self, it, out_dir="."):
        """sxx over the whole grid, i outermost, k innermost; snapshot_1..5
        created empty (ElasticWaveEquation3D.cpp:87-112)."""
        F = self.solver.field("sxx")  # (nz, ny, nx)
        for q in range(1, 6):
            open(os.path.join(out_dir, f"snapshot_{q}_{it}.txt"), "w").close()
        ElasticWaveEquation._write_lines(os.path.join(out_dir, f"snapshot_0_{it}.txt"), F.transpose(2, 1, 0))

    def calculateDamping(self, out_dir="."):
        """The three damping volumes, i outermost (ElasticWaveEquation3D.cpp:
        253-311): d_a(index on axis a) inside the updated box, 0 on the halo."""
        upd = []
        for m in self.n:
            u = np.zeros(m, bool)
            u[5:m - 5] = True
            upd.append(u)
        box = upd[0][:, None, None] & upd[1][None, :, None] & upd[2][None, None, :]
        for a, name in enumerate("XYZ"):
            line = self.damping[a].astype(np.float32).reshape([-1 if b == a else 1 for b in range(3)])
            ElasticWaveEquation._write_lines(os.path.join(out_dir, f"damping{name}.txt"),
                                             np.where(box, line, np.float32(0)))

    def close(self):
        self.solver.close()


def scales_for_maxabs(maxabs, headroom_log2=3, target_log2=14):
    """Power-of-two storage scales that map each field's largest |value| to
    at most 2^(target_log2 - headroom_log2) (fields that stayed 0 get 1)."""
    out = np.ones(9, np.float32)
    for f, m in enumerate(np.asarray(maxabs, np.float64)):
        if m > 0 and np.isfinite(m):
            out[f] = np.float32(2.0 ** (target_log2 - headroom_log2 - int(np.ceil(np.log2(m)))))
    return out
