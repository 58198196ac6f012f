"""Error behaviour of the C ABI that needs no GPU: every configuration check
of sg_create runs before any device call, and every solver entry point
rejects a NULL handle; each failure returns its sg_status (never exits) and
leaves a message in sg_last_error.  (The reference exits through CHECK,
common.h:6-16.)"""
import ctypes

import pytest

SG_OK, SG_ERR_INVALID = 0, -1


@pytest.fixture(scope="module")
def L():
    import sgfd

    sgfd.lib()
    return sgfd.lib()


def _cfg(**kw):
    import sgfd

    c = sgfd._Config(64, 64, 64, 8, 1e-3, 10.0, 10.0, 10.0,
                     (ctypes.c_float * 5)(1.2, -0.09, 0.014, -0.0018, 0.00012), 100, 0, 0, 0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("kw,needle", [
    ({"nx": 10}, b"extent"),
    ({"nz": 0}, b"extent"),
    ({"pml": 28}, b"pml"),
    ({"pml": -1}, b"pml"),
    ({"dt": 0.0}, b"positive"),
    ({"dy": -1.0}, b"positive"),
    ({"time_length": -1}, b"time_length"),
    ({"precision": 4}, b"precision"),
    ({"source_mode": 2}, b"source_mode"),
    ({"kernel": 2}, b"kernel"),
    ({"nx": 32768, "ny": 16384, "nz": 11, "pml": 0}, b"plane"),
])
def test_create_rejects_bad_configurations(L, kw, needle):
    h = ctypes.c_void_p(1)
    assert L.sg_create(ctypes.byref(_cfg(**kw)), ctypes.byref(h)) == SG_ERR_INVALID
    assert h.value is None  # the out handle is cleared
    assert needle in L.sg_last_error()


def test_fp64_runs_the_per_point_kernels_only(L):
    # FP64 (precision 1) with the tiled sweeps (kernel 0) is refused before
    # any device call; FP64 solvers use SG_KERNEL_SIMPLE
    h = ctypes.c_void_p(1)
    assert L.sg_create(ctypes.byref(_cfg(precision=1, kernel=0)), ctypes.byref(h)) == -6  # SG_ERR_UNSUPPORTED
    assert h.value is None
    assert b"FP64" in L.sg_last_error()


def test_tiled_split_arrays_limited_to_4gib_each(L):
    # 16384x16384 plane, P8: each split array holds 2*8*2^28 z-slab slots,
    # 16 GiB in FP32, beyond the tiled sweeps' 32-bit offsets
    h = ctypes.c_void_p(1)
    c = _cfg(nx=16384, ny=16384, nz=64)
    assert L.sg_create(ctypes.byref(c), ctypes.byref(h)) == -6  # SG_ERR_UNSUPPORTED
    assert h.value is None
    assert b"split" in L.sg_last_error()


def test_create_rejects_null_arguments(L):
    h = ctypes.c_void_p()
    assert L.sg_create(None, ctypes.byref(h)) == SG_ERR_INVALID
    assert L.sg_create(ctypes.byref(_cfg()), None) == SG_ERR_INVALID


@pytest.mark.parametrize("name,args", [
    ("sg_step", (1,)),
    ("sg_run", ()),
    ("sg_synchronize", ()),
    ("sg_reset", ()),
    ("sg_set_graph", (1,)),
    ("sg_set_streams", (2,)),
    ("sg_get_streams", (ctypes.byref(ctypes.c_int()),)),
    ("sg_set_rounding", (0,)),
    ("sg_set_compensation", (0,)),
    ("sg_set_source_sample", (0, 1.0)),
    ("sg_set_medium_velocity", (None, None, None)),
])
def test_null_solver_is_rejected(L, name, args):
    fn = getattr(L, name)
    assert fn(None, *args) == SG_ERR_INVALID
    assert b"NULL" in L.sg_last_error()


def test_destroy_null_is_a_no_op(L):
    assert L.sg_destroy(None) == SG_OK


def test_batch_needs_shots(L):
    out = ctypes.c_void_p()
    assert L.sg_batch_create(None, 0, ctypes.byref(out)) == SG_ERR_INVALID


def test_elastic_from_velocity_matches_the_formula_bit_for_bit(L):
    import numpy as np

    import sgfd
    from sgfd import configs

    rng = np.random.default_rng(3)
    vp = (3000 * (1 + 0.1 * rng.standard_normal(5000))).astype(np.float32)
    vs = (vp / np.float32(1.75)).astype(np.float32)
    rho = (310 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    got = sgfd.elastic_from_velocity(vp, vs, rho)
    exp = configs.elastic_from_velocity(vp, vs, rho)
    for g, e in zip(got, exp):
        assert g.dtype == np.float32 and np.array_equal(g, e)


@pytest.mark.parametrize("bad", ["rho0", "vs0", "vpnan"])
def test_elastic_from_velocity_rejects_bad_values(L, bad):
    import numpy as np

    import sgfd

    vp, vs, rho = (np.full(10, v, np.float32) for v in (3000.0, 1700.0, 2000.0))
    {"rho0": rho, "vs0": vs, "vpnan": vp}[bad][7] = np.nan if bad == "vpnan" else 0.0
    with pytest.raises(sgfd.SgError, match="first bad index 7"):
        sgfd.elastic_from_velocity(vp, vs, rho)
