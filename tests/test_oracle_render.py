"""The bounce loop's path-id convention in the oracle (CPU): pid = pixel
rank * spp + sample with pixels ranked in bands of 8 rows, column by column
inside a band.  The C restatement (orc_path_pixel, used by the oracle's
camera) and the numpy inverse (pixel_first_paths, used by its resolve) are
checked against each other and against what the convention fixes: a
bijection, the samples of a pixel adjacent, 64 consecutive single-sample ids
inside one 8x8 tile, and plain row-major order for an image one pixel wide."""
import numpy as np
import pytest

from oracle import orc_py as orc

SIZES = [(1, 1, 1), (1, 7, 1), (7, 3, 2), (8, 8, 1), (13, 17, 3), (16, 12, 4), (9, 16, 5), (40, 23, 1)]


@pytest.mark.parametrize("w,h,spp", SIZES)
def test_path_pixel_is_a_bijection_with_inverse(w, h, spp):
    n = w * h * spp
    xy = orc.path_pixel(np.arange(n), w, h, spp)
    assert xy[:, 0].min() >= 0 and xy[:, 0].max() < w and xy[:, 1].min() >= 0 and xy[:, 1].max() < h
    pix = xy[:, 1] * w + xy[:, 0]
    # each pixel exactly spp times, and its samples are consecutive ids
    assert np.array_equal(np.bincount(pix, minlength=w * h), np.full(w * h, spp))
    first = orc.pixel_first_paths(w, h, spp)
    for s in range(spp):
        assert np.array_equal(pix[first + s], np.arange(w * h))


def test_waves_cover_tiles():
    w, h = 40, 23
    xy = orc.path_pixel(np.arange(w * h), w, h, 1)
    # full bands: each run of 64 ids is an 8x8 tile
    for start in range(0, (h // 8) * 8 * w - 63, 64):
        t = xy[start:start + 64]
        assert np.ptp(t[:, 0]) == 7 and np.ptp(t[:, 1]) == 7
    # the last band (h % 8 = 7 rows) walks columns of 7
    last = xy[(h // 8) * 8 * w:]
    assert np.array_equal(last[:7, 0], np.zeros(7)) and np.array_equal(last[:7, 1], np.arange(16, 23))


def test_single_column_is_row_major():
    xy = orc.path_pixel(np.arange(11), 1, 11, 1)
    assert np.array_equal(xy[:, 0], np.zeros(11)) and np.array_equal(xy[:, 1], np.arange(11))


def test_resolve_averages_each_pixels_samples_in_order():
    w, h, spp = 9, 10, 3
    n = w * h * spp
    L = np.zeros((n, 4), np.float32)
    xy = orc.path_pixel(np.arange(n), w, h, spp)
    L[:, 0] = xy[:, 1] * w + xy[:, 0]          # pixel index in every sample
    L[:, 1] = np.arange(n) % spp               # sample index
    img = orc.resolve(L, w, h, spp)
    np.testing.assert_array_equal(img[:, 0], np.arange(w * h, dtype=np.float32))
    np.testing.assert_allclose(img[:, 1], np.float32(1.0), rtol=1e-7)
