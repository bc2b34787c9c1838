"""The transcendental functions shared by the GPU kernels and the oracle
(mambatracer_amd/csrc/mt_xmath.h: sin, cos, acos, atan2, x^5), pinned against
float64 numpy rounded to float32 over dense sweeps of the ranges the path
tracer uses (angles in [-8, 8] and up to 64, cosines in [-1, 1], direction
components) plus the special values C99 fixes (signed zeros, infinities,
NaN).  Bound: the OpenCL 1.2 full-profile accuracy the reference's builtins
are held to (sin, cos, acos 4 ulp; atan2 6 ulp; pow 16 ulp), and at least
half the results correctly rounded; the errors measured are 1-3 ulp."""
import numpy as np
import pytest

from oracle import orc_py as orc


def ulps(got, want):
    g = got.astype(np.float32).view(np.int32).astype(np.int64)
    w = want.astype(np.float32).view(np.int32).astype(np.int64)
    g = np.where(g < 0, np.int64(-2**31) - g, g)   # monotone integer order of floats
    w = np.where(w < 0, np.int64(-2**31) - w, w)
    return np.abs(g - w)


def check(got, want64, bound=4, frac_exact=0.45):
    want = want64.astype(np.float32)
    u = ulps(got, want)
    assert u.max() <= bound, f"max error {u.max()} ulp at {np.argmax(u)}"
    assert (u == 0).mean() >= frac_exact, f"only {(u == 0).mean():.5f} correctly rounded"


def test_sin_cos_dense():
    x = np.linspace(-8.0, 8.0, 2_000_001, dtype=np.float32)
    x = np.concatenate([x, np.float32(np.pi) * np.arange(-5, 6, dtype=np.float32) / 4,
                        np.float32(3.14159265359) * np.arange(0, 3, dtype=np.float32)])
    check(orc.xm_eval(orc.XM_SIN, x), np.sin(x.astype(np.float64)))
    check(orc.xm_eval(orc.XM_COS, x), np.cos(x.astype(np.float64)))


def test_sin_cos_small_and_random():
    rng = np.random.default_rng(5)
    x = np.concatenate([np.float32(10.0) ** rng.uniform(-30, 0.9, 200_000).astype(np.float32),
                        rng.uniform(-64, 64, 200_000).astype(np.float32)])
    x = np.concatenate([x, -x])
    check(orc.xm_eval(orc.XM_SIN, x), np.sin(x.astype(np.float64)))
    check(orc.xm_eval(orc.XM_COS, x), np.cos(x.astype(np.float64)))
    # odd / even symmetry exactly
    assert np.array_equal(orc.xm_eval(orc.XM_SIN, -x), -orc.xm_eval(orc.XM_SIN, x))
    assert np.array_equal(orc.xm_eval(orc.XM_COS, -x), orc.xm_eval(orc.XM_COS, x))


def test_acos_dense_and_ends():
    y = np.linspace(-1.0, 1.0, 2_000_001, dtype=np.float32)
    ends = np.nextafter(np.float32(1), np.float32(0)) - np.arange(64, dtype=np.float32) * np.float32(6e-8)
    y = np.concatenate([y, ends, -ends])
    check(orc.xm_eval(orc.XM_ACOS, y), np.arccos(y.astype(np.float64)))
    a = orc.xm_eval(orc.XM_ACOS, np.array([1, -1, 0, 2, np.nan], np.float32))
    assert a[0] == 0 and not np.signbit(a[0])
    assert a[1] == np.float32(np.pi) and a[2] == np.float32(np.pi / 2)
    assert np.isnan(a[3]) and np.isnan(a[4])


def test_atan2_directions():
    rng = np.random.default_rng(6)
    d = rng.normal(size=(1_000_000, 2)).astype(np.float32)
    d = np.concatenate([d, d * np.float32(1e-20), d * np.float32(1e20)])
    y, x = d[:, 0], d[:, 1]
    check(orc.xm_eval(orc.XM_ATAN2, y, x), np.arctan2(y.astype(np.float64), x.astype(np.float64)), bound=6)
    # near the axes and the diagonals, where the argument reduction switches
    t = np.float32(1) + np.arange(-2000, 2001, dtype=np.float32) * np.float32(1e-7)
    for yy, xx in [(t, np.ones_like(t)), (np.full_like(t, 0.2679491924), t),
                   (np.arange(-1000, 1001, dtype=np.float32) * np.float32(1e-6), -np.ones(2001, np.float32))]:
        check(orc.xm_eval(orc.XM_ATAN2, yy, xx), np.arctan2(yy.astype(np.float64), xx.astype(np.float64)), bound=6,
              frac_exact=0.3)


@pytest.mark.parametrize("y,x", [(0.0, 0.0), (-0.0, 0.0), (0.0, -0.0), (-0.0, -0.0), (1.0, 0.0), (1.0, -0.0),
                                 (-1.0, 0.0), (0.0, -1.0), (-0.0, -1.0), (np.inf, np.inf), (-np.inf, np.inf),
                                 (np.inf, -np.inf), (np.inf, 1.0), (1.0, np.inf), (1.0, -np.inf),
                                 (-1.0, -np.inf)])
def test_atan2_special_values(y, x):
    got = orc.xm_eval(orc.XM_ATAN2, np.array([y], np.float32), np.array([x], np.float32))[0]
    want = np.float32(np.arctan2(np.float64(y), np.float64(x)))
    assert got == want and np.signbit(got) == np.signbit(want)


def test_atan2_nan():
    a = orc.xm_eval(orc.XM_ATAN2, np.array([np.nan, 1.0], np.float32), np.array([1.0, np.nan], np.float32))
    assert np.isnan(a).all()


def test_pow5():
    x = np.linspace(0.0, 1.0, 1_000_001, dtype=np.float32)
    check(orc.xm_eval(orc.XM_POW5, x), x.astype(np.float64) ** 5, bound=16, frac_exact=0.4)
