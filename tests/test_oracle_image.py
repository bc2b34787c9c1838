"""Oracle checks for the image pipeline (CImage.processImage) and the float
scan (PrefixSumFloat), pinned by known answers and closed forms rather than
by retyping the oracle's formulas."""
import math

import numpy as np

from oracle import orc_py as orc


def _i32(u):
    return np.array(u, dtype=np.uint32).astype(np.int32).item()


def test_argb_known_answers():
    # getIntARGB (Util.cl:469-480): clamp, * 255, truncate, pack A R G B
    assert orc.argb(1.0, 0.5, 0.0, 1.0) == _i32(0xFFFF7F00)
    assert orc.argb(2.0, -1.0, float("nan"), 1.0) == _i32(0xFFFF0000)
    assert orc.argb(0.999, 0.004, 0.0039, 0.5) == _i32(0x7FFE0100)  # 254, 1, 0
    assert orc.image_init_argb() == _i32(0xFF000000)


def _gray_expected(v, exposure, gamma, frame_count):
    """updateFrameImage on a uniform gray frame, in float64 from the
    published Reinhard / xyY formulas (RGBSpace.cl:23-106)."""
    c = v / frame_count
    lum = 0.2989 * c + 0.5866 * c + 0.1145 * c
    avg = 0.01 + lum                      # exp(mean(log(0.01 + lum)))
    s = exposure * lum / avg
    Y = s / (1 + s)
    X = (0.5893 + 0.1789 + 0.1831) * c
    Yc = (0.2904 + 0.6051 + 0.1045) * c
    Z = (0.0 + 0.0684 + 1.0202) * c
    x, y = X / (X + Yc + Z), Yc / (X + Yc + Z)
    X2, Z2 = Y * x / y, Y * (1 - x - y) / y
    rgb = np.array([1.967 * X2 - 0.548 * Y - 0.297 * Z2,
                    -0.955 * X2 + 1.938 * Y - 0.027 * Z2,
                    0.064 * X2 - 0.130 * Y + 0.982 * Z2])
    return rgb ** (1 / gamma), math.log(0.01 + lum)


def test_gray_frame_closed_form():
    n = 5000
    accum = np.zeros((n, 4), np.float32)
    accum[:, :3] = 0.9
    frame, argb, (tl, tc) = orc.image_process(accum, 3.0, 0.18, 2.2)
    rgb, lg = _gray_expected(0.9, 0.18, 2.2, 3.0)
    assert tc == n
    assert abs(tl - n * lg) <= 1e-5 * abs(n * lg)
    np.testing.assert_allclose(frame[:, :3], np.broadcast_to(rgb, (n, 3)), rtol=2e-5)
    assert (frame[:, 3] == 1).all()
    want = (0xFF << 24) | (int(min(max(rgb[0], 0), 1) * 255) << 16) | (int(min(max(rgb[1], 0), 1) * 255) << 8) \
        | int(min(max(rgb[2], 0), 1) * 255)
    assert (argb == _i32(want)).all()


def test_black_pixels_keep_argb_and_totals_reuse():
    n = 3000
    rng = np.random.default_rng(1)
    accum = np.zeros((n, 4), np.float32)
    accum[::3, :3] = rng.uniform(0.1, 5.0, size=(len(accum[::3]), 3))
    sentinel = np.full(n, 12345, np.int32)
    frame, argb, (tl, tc) = orc.image_process(accum, 1.0, 0.18, 2.2, argb=sentinel)
    assert tc == len(accum[::3])
    black = np.ones(n, bool)
    black[::3] = False
    assert (argb[black] == 12345).all() and (argb[~black] != 12345).all()
    np.testing.assert_array_equal(frame[black, :3], 0.0)
    # totals: sum of log(0.01 + Lw) over lit pixels, float64 reference
    lw = (0.2989 * accum[:, 0].astype(np.float64) + 0.5866 * accum[:, 1] + 0.1145 * accum[:, 2])
    want = np.log(0.01 + lw[~black]).sum()
    assert abs(tl - want) <= 1e-5 * abs(want)
    # reusing totals (frameCount >= 3) reproduces the same frame
    frame2, argb2, t2 = orc.image_process(accum, 1.0, 0.18, 2.2, argb=sentinel, totals=(tl, tc))
    assert t2 == (tl, tc)
    np.testing.assert_array_equal(frame2, frame)
    np.testing.assert_array_equal(argb2, argb)


def test_float_scan_matches_float64_prefix_sums():
    rng = np.random.default_rng(2)
    for n in [1, 7, 2047, 2048, 2049, 300_001]:
        x = rng.uniform(-1.0, 2.0, size=n).astype(np.float32)
        out, tot = orc.scan_excl_f32(x)
        exact = np.concatenate([[0.0], np.cumsum(x.astype(np.float64))])
        bound = 1e-6 * np.concatenate([[0.0], np.cumsum(np.abs(x.astype(np.float64)))]) + 1e-30
        assert (np.abs(out - exact[:-1]) <= 8 * bound[:-1] + 1e-7).all(), n
        assert abs(tot - exact[-1]) <= 8 * bound[-1] + 1e-7
        assert out[0] == 0.0


def test_float_scan_integers_exact_and_order_fixed():
    # small integers: every partial sum is exact in float, so any order agrees
    x = np.arange(1, 70_001, dtype=np.float32) % 7
    out, tot = orc.scan_excl_f32(x)
    exact = np.concatenate([[0.0], np.cumsum(x.astype(np.float64))])
    np.testing.assert_array_equal(out, exact[:-1].astype(np.float32))
    assert tot == exact[-1]
    # a case whose rounding depends on the order: 1e8 then ones; the fixed
    # order gives the same answer every call
    y = np.ones(10_000, np.float32)
    y[0] = 1e8
    a = orc.scan_excl_f32(y)
    b = orc.scan_excl_f32(y)
    np.testing.assert_array_equal(a[0], b[0])
    assert a[1] == b[1]
