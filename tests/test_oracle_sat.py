"""Pins the oracle's stage C without the reference: the reference's own SAT
known answers (test/lightgrid/TestSimpleGrid.java:39-67,
TestLightGrid.java:41-62), float64 summed-area tables, and sampling against
an independent float64 pbrt piecewise-constant Distribution2D
(test/sampling/Distribution1D.java / Distribution2D.java restated below in
numpy) on the same function."""
import math

import numpy as np
import pytest

from oracle import orc_py as orc
from tests.conftest import golden


# ---------------------------------------------------------------- SAT build
def test_simplegrid_impulse_known_answer():
    k = golden("simplegrid_impulse.json")
    f = np.array(k["func_rows"], np.float32)
    sat = orc.sat_tiles(f.reshape(-1), k["nu"], k["nv"], k["rx"], k["ry"]).reshape(16, 16)
    np.testing.assert_array_equal(sat, np.array(k["sat_rows"], np.float32))


def test_lightgrid_all_ones_known_answer():
    k = golden("lightgrid_all_ones.json")
    func = orc.lg_set_tile_func(np.full(512, k["tile_value"], np.float32))
    assert (func == 1).all()
    sat = orc.sat_tiles(func, k["nu"], k["nv"], k["rx"], k["ry"]).reshape(1600, 1600)
    np.testing.assert_array_equal(sat, np.tile(np.array(k["tile_sat_rows"], np.float32), (50, 100)))


def test_tile_sat_matches_float64():
    rng = np.random.default_rng(1)
    f = rng.random((64, 64)).astype(np.float32)
    sat = orc.sat_tiles(f.reshape(-1), 64, 64, 16, 32).reshape(64, 64)
    ref = np.zeros((64, 64))
    for ty in range(2):
        for tx in range(4):
            blk = f[32 * ty:32 * ty + 32, 16 * tx:16 * tx + 16].astype(np.float64)
            ref[32 * ty:32 * ty + 32, 16 * tx:16 * tx + 16] = blk.cumsum(1).cumsum(0)
    np.testing.assert_allclose(sat, ref, rtol=1e-5)


def test_kogge_stone_tree_order_differs_from_serial():
    """The tile scan is the Kogge-Stone tree, not a serial sum: on inputs
    where the two orders round differently the oracle follows the tree."""
    e = 2.0 ** -24  # half an ulp of 1.0: serial adds round away, the tree keeps two of them
    x = np.array([1.0, e, e, e] + [0.0] * 12, np.float32)
    sat = orc.sat_tiles(x, 16, 1, 16, 1)
    serial = np.float32(0)
    for v in x[:4]:
        serial = np.float32(serial + v)
    tree = np.float32(np.float32(x[0] + x[1]) + np.float32(x[2] + x[3]))
    assert sat[3] == tree and sat[3] != serial


def test_env_sat_matches_float64_and_serial_order():
    rng = np.random.default_rng(2)
    w, h = 97, 41
    rgba = rng.random((h * w, 4)).astype(np.float32)
    lum, sat = orc.env_lum_sat(rgba, w, h)
    L = (0.2989 * rgba[:, 0].astype(np.float64) + 0.5866 * rgba[:, 1] + 0.1145 * rgba[:, 2]).reshape(h, w)
    np.testing.assert_allclose(lum.reshape(h, w), L, rtol=1e-6)
    np.testing.assert_allclose(sat.reshape(h, w), L.cumsum(1).cumsum(0), rtol=1e-5)
    # rows are summed left to right in float32 before the columns (calculateSAT order)
    row0 = np.float32(0)
    for v in lum[:w]:
        row0 = np.float32(row0 + v)
    assert sat[w - 1] == row0


# ---------------------------------------------------------------- Distribution2D
class Dist1D:
    """test/sampling/Distribution1D.java in float64."""

    def __init__(self, f):
        self.f = np.asarray(f, np.float64)
        n = len(f)
        self.cdf = np.concatenate([[0.0], np.cumsum(self.f / n)])
        self.integral = self.cdf[-1]
        self.cdf = self.cdf / self.integral if self.integral > 0 else np.arange(n + 1) / n

    def sample(self, u):
        n = len(self.f)
        i = int(np.searchsorted(self.cdf[:n], u, side="right"))
        off = max(0, i - 1)
        du = (u - self.cdf[off]) / (self.cdf[off + 1] - self.cdf[off])
        return (off + du) / n, off, self.f[off] / self.integral


class Dist2D:
    """test/sampling/Distribution2D.java in float64."""

    def __init__(self, f):
        self.rows = [Dist1D(r) for r in f]
        self.marg = Dist1D([r.integral for r in self.rows])

    def sample(self, u0, u1):
        v, oy, pv = self.marg.sample(u1)
        u, ox, pu = self.rows[oy].sample(u0)
        return (u, v), (ox, oy), pu * pv


def _near_boundary(d2, u0, u1, off, eps=1e-5):
    ox, oy = off
    m = d2.marg.cdf
    c = d2.rows[oy].cdf
    return min(abs(u1 - m[oy]), abs(u1 - m[oy + 1]), abs(u0 - c[ox]), abs(u0 - c[ox + 1])) < eps


@pytest.mark.parametrize("region", [(0, 0, 48, 32), (16, 8, 16, 16), (5, 3, 20, 11)])
def test_region_sampling_against_distribution2d(region):
    rng = np.random.default_rng(3)
    nu, nv = 48, 32
    func = (rng.random((nv, nu)) ** 2 + 0.05).astype(np.float32)
    lum, sat = orc.env_lum_sat(np.repeat(func.reshape(-1, 1), 4, 1), nu, nv)
    mx, my, sx, sy = region
    sub = lum.reshape(nv, nu)[my:my + sy, mx:mx + sx]
    d2 = Dist2D(sub)
    u = rng.random((4000, 2)).astype(np.float32)
    uv, off, pdf = orc.sat_sample_region(lum, sat, nu, nv, mx, my, sx, sy, u)
    checked = 0
    for k in range(len(u)):
        (ru, rv), (ox, oy), rp = d2.sample(float(u[k, 0]), float(u[k, 1]))
        if _near_boundary(d2, float(u[k, 0]), float(u[k, 1]), (ox, oy)):
            continue
        checked += 1
        assert off[k, 0] == ox + mx and off[k, 1] == oy + my
        # absolute uv: (offset + du) / full map size
        np.testing.assert_allclose(uv[k], [(ru * sx + mx) / nu, (rv * sy + my) / nv], atol=2e-4)
        np.testing.assert_allclose(pdf[k], rp, rtol=2e-3)
    assert checked > 3900


def test_sampling_1d_reference_case():
    """test/sampling/Test.java:17: {0,1,0,1,1,1,0,0}; zero-valued cells are never picked."""
    k = golden("distribution1d_case.json")
    f = np.array(k["func"], np.float32)
    # the fixture's closed-form samples
    us = np.array([[c["u"], 0.5] for c in k["samples"]], np.float32)
    kuv, koff, kpdf = orc.sat_sample_region(*orc.env_lum_sat(np.repeat(f.reshape(-1, 1), 4, 1), 8, 1), 8, 1, 0, 0,
                                            8, 1, us)
    for j, c in enumerate(k["samples"]):
        assert koff[j, 0] == c["offset"]
        assert abs(kuv[j, 0] - c["u_out"]) < 1e-6 and abs(kpdf[j] - c["pdf"]) < 1e-5
    lum, sat = orc.env_lum_sat(np.repeat(f.reshape(-1, 1), 4, 1), 8, 1)
    u = np.stack([np.linspace(0, 0.999, 400), np.full(400, 0.5)], 1).astype(np.float32)
    uv, off, pdf = orc.sat_sample_region(lum, sat, 8, 1, 0, 0, 8, 1, u)
    assert set(off[:, 0].tolist()) == {1, 3, 4, 5}
    d = Dist1D(f)
    for k in range(400):
        x, o, p = d.sample(float(u[k, 0]))
        if min(abs(u[k, 0] - d.cdf[o]), abs(u[k, 0] - d.cdf[o + 1])) < 1e-6:
            continue
        assert off[k, 0] == o
        assert abs(uv[k, 0] - x) < 1e-6
        assert abs(pdf[k] - p) < 1e-5       # = 8/4 = 2 on picked cells
    np.testing.assert_allclose(pdf, 2.0, rtol=1e-6)


def test_lightgrid_sampling_against_distribution2d():
    rng = np.random.default_rng(4)
    accum = (rng.random((1600, 1600)) ** 3 + 0.01).astype(np.float32)
    sat = orc.sat_tiles(accum.reshape(-1), 1600, 1600, 16, 32)
    cells = rng.integers(0, 5000, size=300).astype(np.int32)
    u = rng.random((300, 2)).astype(np.float32)
    txy, uv, pdf = orc.lg_sample(accum.reshape(-1), sat, cells, u)
    checked = 0
    for k in range(300):
        s = int(cells[k])
        cx, cy = s % 100, s // 100
        tile = accum[32 * cy:32 * cy + 32, 16 * cx:16 * cx + 16]
        d2 = Dist2D(tile)
        (ru, rv), (ox, oy), rp = d2.sample(float(u[k, 0]), float(u[k, 1]))
        if _near_boundary(d2, float(u[k, 0]), float(u[k, 1]), (ox, oy)):
            continue
        checked += 1
        assert txy[k, 0] == ox and txy[k, 1] == oy
        np.testing.assert_allclose(uv[k], [ru, rv], atol=1e-4)
        np.testing.assert_allclose(pdf[k], rp, rtol=1e-3)
    assert checked > 290


def test_uniform_environment_pdf_closed_form():
    """Constant luminance: the continuous SAT pdf is 1, so the solid-angle pdf
    is 1 / (2 pi^2 sin(theta)) and every direction is equally likely per area."""
    w, h = 64, 32
    rgba = np.ones((w * h, 4), np.float32)
    lum, sat = orc.env_lum_sat(rgba, w, h)
    u = np.random.default_rng(5).random((2000, 2)).astype(np.float32)
    d, c, pw, tx = orc.env_illuminate(rgba, lum, sat, w, h, u)
    D = d[:, :3].astype(np.float64)
    sin_theta = np.hypot(D[:, 0], D[:, 2]) / np.linalg.norm(D, axis=1)   # well conditioned at the poles
    ok = sin_theta > 1e-3
    np.testing.assert_allclose(pw[ok], 1.0 / (2 * math.pi ** 2 * sin_theta[ok]), rtol=2e-3)
    np.testing.assert_allclose(np.linalg.norm(d[:, :3], axis=1), 1.0, atol=1e-6)
    c2, pw2, tx2 = orc.env_radiance(rgba, lum, sat, w, h, d)
    assert (c2[:, :3] == 1).all()
    # getRadianceEnvironmentLight evaluates sin at the texel row's top edge v = y / h
    # (Light.cl:281-282); row 0 divides by sin(0) = 0, as in the reference
    # (theta = -pi maps to column x = w, outside the map: pdf 0 and the texel
    # index wraps to the next row, Util.cl:405,416 -- excluded here)
    y = tx2 // w
    wrap = pw2 == 0
    assert wrap.mean() < 0.01
    top = (y > 0) & ~wrap
    np.testing.assert_allclose(pw2[top], 1.0 / (2 * math.pi ** 2 * np.sin(math.pi * y[top] / h)), rtol=1e-5)
    assert np.isinf(pw2[(y == 0) & ~wrap]).all()


# ---------------------------------------------------------------- adaptive environment light
def _sph_bin(d, w, h):
    """getSphericalGridXY (Util.cl:397-410) in float64: the (x, y) bin of direction d."""
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    phi = np.arccos(np.clip(d[:, 1], -1, 1))
    theta = np.arctan2(d[:, 2], d[:, 0])
    return ((0.5 - 0.5 * theta / math.pi) * w).astype(np.int64), (phi / math.pi * h).astype(np.int64)


def _adaptive_inputs(rng, n, w, h, uniform):
    rgba = np.ones((w * h, 4), np.float32) if uniform else \
        np.concatenate([(rng.random((w * h, 3)) ** 2 + 0.01), np.ones((w * h, 1))], 1).astype(np.float32)
    lum, sat = orc.env_lum_sat(rgba, w, h)
    accum = (rng.random(1600 * 1600) ** 3 + 0.01).astype(np.float32)
    lg_sat = orc.sat_tiles(accum, 1600, 1600, 16, 32)
    cam = np.array([0.3, -0.2, 0.1], np.float32)
    hit = np.concatenate([rng.uniform(-3, 3, (n, 3)), np.zeros((n, 1))], 1).astype(np.float32)
    u = rng.random((n, 4)).astype(np.float32)
    return rgba, lum, sat, accum, lg_sat, cam, hit, u


@pytest.mark.parametrize("uniform", [False, True])
def test_adaptive_illuminate_structure(uniform):
    """illuminateEnvironmentLight1 pinned by its structure: the accum entry
    names the camera cell (independent float64 spherical binning) and the
    tile the cell's distribution picks (lg_sample, pinned above); the texel
    lies in that tile's rectangle of the map (closed form: tile (tx, ty) of a
    w x h map covers x in [tx w/16, (tx+1) w/16), y in [ty h/32, (ty+1)
    h/32)) and equals region sampling of that rectangle; on a constant map
    the pdf is 1 / (2 pi^2 sin(theta))."""
    rng = np.random.default_rng(21)
    n, w, h = 4000, 256, 128
    rgba, lum, sat, accum, lg_sat, cam, hit, u = _adaptive_inputs(rng, n, w, h, uniform)
    d, c, pw, lo, gi, txy = orc.env_illuminate_adaptive(rgba, lum, sat, w, h, accum, lg_sat, cam, hit, u)
    gx, gy = gi % 1600, gi // 1600
    cx, lx, cy, ly = gx // 16, gx % 16, gy // 32, gy % 32
    bx, by = _sph_bin(hit[:, :3].astype(np.float64) - cam.astype(np.float64), 100, 50)
    assert ((bx == cx) & (by == cy)).mean() > 0.999
    tl, _, _ = orc.lg_sample(accum, lg_sat, (cx + 100 * cy).astype(np.int32), u[:, :2])
    np.testing.assert_array_equal(tl[:, 0], lx)
    np.testing.assert_array_equal(tl[:, 1], ly)
    sx, sy = w // 16, h // 32
    assert ((txy[:, 0] >= lx * sx) & (txy[:, 0] < (lx + 1) * sx)).all()
    assert ((txy[:, 1] >= ly * sy) & (txy[:, 1] < (ly + 1) * sy)).all()
    for k in range(0, n, 97):
        _, off, pdf = orc.sat_sample_region(lum, sat, w, h, int(lx[k]) * sx, int(ly[k]) * sy, sx, sy, u[k:k + 1, 2:])
        assert tuple(off[0]) == tuple(txy[k])
    np.testing.assert_array_equal(lo, np.array([orc_lum(x) for x in c], np.float32))
    if uniform:
        D = d[:, :3].astype(np.float64)
        sin_theta = np.hypot(D[:, 0], D[:, 2])
        ok = sin_theta > 1e-3
        np.testing.assert_allclose(pw[ok], 1.0 / (2 * math.pi ** 2 * sin_theta[ok]), rtol=2e-3)


def orc_lum(c):
    """Luminance, RGBSpace.cl:28-32, in float32 left to right."""
    r, g, b = (np.float32(x) for x in c[:3])
    return np.float32(np.float32(np.float32(0.2989) * r + np.float32(0.5866) * g) + np.float32(0.1145) * b)


def test_adaptive_radiance_structure_and_accumulation():
    """getRadianceEnvironmentLight1 + accumLightGrid: the entry is (camera
    cell, tile of the ray direction) by independent float64 binning; on a
    constant map the in-tile pdf is 1, so pdfa = 1 / (2 pi^2 sin(pi y / h))
    for texels inside their tile; accum gains count * luminance per entry."""
    rng = np.random.default_rng(22)
    n, w, h = 20000, 256, 128
    rgba, lum, sat, _, _, cam, hit, _ = _adaptive_inputs(rng, n, w, h, True)
    dirs = rng.normal(size=(n, 3))
    dirs = np.concatenate([dirs / np.linalg.norm(dirs, axis=1, keepdims=True), np.zeros((n, 1))], 1).astype(np.float32)
    acc = np.zeros(1600 * 1600, np.float32)
    c, pa, lo, gi = orc.env_radiance_adaptive(rgba, lum, sat, w, h, cam, dirs, hit, accum=acc)
    gx, gy = gi % 1600, gi // 1600
    bx, by = _sph_bin(hit[:, :3].astype(np.float64) - cam.astype(np.float64), 100, 50)
    tx, ty = _sph_bin(dirs[:, :3].astype(np.float64), 16, 32)
    same = (bx == gx // 16) & (by == gy // 32) & (tx == gx % 16) & (ty == gy % 32)
    assert same.mean() > 0.999
    x, y = _sph_bin(dirs[:, :3].astype(np.float64), w, h)
    inside = (x // (w // 16) == gx % 16) & (y // (h // 32) == gy % 32) & (y > 0) & (x < w)
    assert inside.mean() > 0.99
    np.testing.assert_allclose(pa[inside], 1.0 / (2 * math.pi ** 2 * np.sin(math.pi * y[inside] / h)), rtol=1e-5)
    L = float(lo[0])
    assert (lo == lo[0]).all()
    cnt = np.bincount(gi, minlength=1600 * 1600)
    np.testing.assert_allclose(acc, cnt * L, rtol=1e-6)
    assert cnt.max() >= 2
