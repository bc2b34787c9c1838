"""Stage C on the GPU against the oracle: light-grid tile SAT build and
sampling, the reference's SAT known answers, environment luminance/SAT,
environment sampling and the light-grid guided environment light, all
bit-exact (both sides run the same IEEE operations, the transcendental
functions included: mt_xmath.h).  Any mismatch is listed."""
import numpy as np
import pytest
import torch

from mambatracer_amd import api
from oracle import orc_py as orc
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def test_simplegrid_impulse_known_answer(ctx):
    """test/lightgrid/TestSimpleGrid.java:39-67: 16x16 grid, 0.01 at row 11 col 10."""
    k = golden("simplegrid_impulse.json")
    f = np.array(k["func_rows"], np.float32)
    sat = api.sat_tiles(ctx, ctx.to_device(f.reshape(-1)), k["nu"], k["nv"], k["rx"], k["ry"]).cpu().numpy()
    np.testing.assert_array_equal(sat.reshape(16, 16), np.array(k["sat_rows"], np.float32))


def test_lightgrid_all_ones_known_answer(ctx):
    """test/lightgrid/TestLightGrid.java:41-62: all-ones tile -> sat = (x+1)(y+1) per tile."""
    k = golden("lightgrid_all_ones.json")
    lg = api.CAdaptiveEnvironment(ctx)
    lg.setHDRLuminance(np.full(512, k["tile_value"], np.float32))
    _, func, sat = lg.read_arrays()
    exp = np.tile(np.array(k["tile_sat_rows"], np.float32), (50, 100))
    np.testing.assert_array_equal(func, np.ones_like(func))
    np.testing.assert_array_equal(sat, exp)


def test_lightgrid_update_bit_exact(ctx):
    print("[case] lightgrid update", flush=True)
    rng = np.random.default_rng(1)
    accum = (rng.random((1600, 1600)) ** 4 * 100).astype(np.float32)
    lg = api.CAdaptiveEnvironment(ctx)
    lg.write_accum(accum)
    lg.update()
    a, func, sat = lg.read_arrays()
    np.testing.assert_array_equal(func, accum)
    osat = orc.sat_tiles(accum.reshape(-1), 1600, 1600, 16, 32).reshape(1600, 1600)
    np.testing.assert_array_equal(sat.view(np.int32), osat.view(np.int32))


def test_lightgrid_sampling_bit_exact(ctx):
    print("[case] lightgrid sampling", flush=True)
    rng = np.random.default_rng(2)
    accum = (rng.random((1600, 1600)) ** 3).astype(np.float32)
    accum[rng.random((1600, 1600)) < 0.2] = 0.0
    lg = api.CAdaptiveEnvironment(ctx)
    lg.write_accum(accum)
    lg.update()
    n = 200000
    cell = rng.integers(0, 5000, size=n).astype(np.int32)
    u = rng.random((n, 2)).astype(np.float32)
    txy, uv, pdf = lg.sample(ctx.to_device(cell), ctx.to_device(u))
    osat = orc.sat_tiles(accum.reshape(-1), 1600, 1600, 16, 32)
    otxy, ouv, opdf = orc.lg_sample(accum.reshape(-1), osat, cell, u)
    np.testing.assert_array_equal(txy.cpu().numpy(), otxy)
    np.testing.assert_array_equal(uv.cpu().numpy().view(np.int32), ouv.view(np.int32))
    np.testing.assert_array_equal(pdf.cpu().numpy().view(np.int32), opdf.view(np.int32))
    # the per-cell tables (used above) and the SAT corners give the same bits:
    # handing out the raw arrays switches sampling to the SAT until the next build
    f = api.C.c_void_p()
    api.call("mt_lightgrid_arrays", lg.handle, None, api.C.byref(f), None)
    txy2, uv2, pdf2 = lg.sample(ctx.to_device(cell), ctx.to_device(u))
    np.testing.assert_array_equal(txy2.cpu().numpy(), otxy)
    np.testing.assert_array_equal(uv2.cpu().numpy().view(np.int32), ouv.view(np.int32))
    np.testing.assert_array_equal(pdf2.cpu().numpy().view(np.int32), opdf.view(np.int32))


def bits(name, a, b):
    """Entries whose bits differ (NaN matches NaN); prints the first few."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float32:
        bad = (a.view(np.int32) != b.view(np.int32)) & ~(np.isnan(a) & np.isnan(b))
    else:
        bad = a != b
    if bad.any():
        idx = np.argwhere(bad)[:6]
        print(f"{name}: {int(bad.sum())} differ, e.g. {idx.tolist()} {a[tuple(idx.T)].tolist()} vs "
              f"{b[tuple(idx.T)].tolist()}", flush=True)
    return int(bad.sum())


def test_lightgrid_cells(ctx):
    """Direction -> cell / tile (acos / atan2 binning), exact."""
    rng = np.random.default_rng(3)
    n = 100000
    v = rng.normal(size=(n, 4))
    v[:, 3] = 0
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)  # ray directions are unit
    cell, txy = api.lightgrid_cells(ctx, ctx.to_device(v))
    ocell, otxy = orc.lg_cells(v)
    assert bits("cell", cell.cpu().numpy(), ocell) == 0
    assert bits("tile", txy.cpu().numpy(), otxy) == 0


def _env_map(w=256, h=128):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    sky = 0.2 + 0.8 * (1.0 - (y + 0.5) / h)
    rgba = np.zeros((h, w, 4), np.float32)
    rgba[..., 0] = sky * 0.7
    rgba[..., 1] = sky * 0.85
    rgba[..., 2] = sky
    d = np.hypot((x - 60) / 2.0, (y - 30) / 2.0)
    rgba[..., :3] += (1000.0 * np.exp(-0.5 * d * d))[..., None].astype(np.float32)
    return rgba


@pytest.mark.parametrize("size", [(256, 128), (1024, 512), (100, 37)])
def test_env_lum_sat_bit_exact(ctx, size):
    w, h = size
    rgba = _env_map(w, h)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    olum, osat = orc.env_lum_sat(rgba.reshape(-1, 4), w, h)
    np.testing.assert_array_equal(env.lum.cpu().numpy().view(np.int32), olum.view(np.int32))
    np.testing.assert_array_equal(env.sat.cpu().numpy().view(np.int32), osat.view(np.int32))


def test_env_illuminate_and_radiance(ctx):
    print("[case] env sampling", flush=True)
    w, h = 1024, 512
    rgba = _env_map(w, h)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    rng = np.random.default_rng(4)
    n = 100000
    u = rng.random((n, 2)).astype(np.float32)
    d, c, pw, tx = env.illuminate(ctx.to_device(u))
    olum, osat = orc.env_lum_sat(rgba.reshape(-1, 4), w, h)
    od, oc, opw, otx = orc.env_illuminate(rgba.reshape(-1, 4), olum, osat, w, h, u)
    d, pw, tx = d.cpu().numpy(), pw.cpu().numpy(), tx.cpu().numpy()
    assert bits("dir", d, od) + bits("texel", tx, otx) + bits("pdf", pw, opw) + bits("contrib", c.cpu().numpy(), oc) == 0
    # radiance at the sampled directions
    c2, pw2, tx2 = env.radiance(ctx.to_device(od))
    oc2, opw2, otx2 = orc.env_radiance(rgba.reshape(-1, 4), olum, osat, w, h, od)
    assert bits("texel2", tx2.cpu().numpy(), otx2) + bits("pdf2", pw2.cpu().numpy(), opw2) == 0


def test_region_sampling_bit_exact(ctx):
    rng = np.random.default_rng(6)
    w, h = 64, 48
    func = (rng.random(w * h) ** 2).astype(np.float32)
    _, sat = orc.env_lum_sat(np.stack([func, func, func, func], 1), w, h)
    lum = orc.env_lum_sat(np.stack([func, func, func, func], 1), w, h)[0]
    u = rng.random((5000, 2)).astype(np.float32)
    for (mx, my, sx, sy) in [(0, 0, w, h), (16, 8, 16, 16), (40, 30, 24, 18)]:
        uv, off, pdf = api.sat_sample_region(ctx, ctx.to_device(lum), ctx.to_device(sat), w, h, mx, my, sx, sy,
                                             ctx.to_device(u))
        ouv, ooff, opdf = orc.sat_sample_region(lum, sat, w, h, mx, my, sx, sy, u)
        np.testing.assert_array_equal(off.cpu().numpy(), ooff)
        np.testing.assert_array_equal(uv.cpu().numpy().view(np.int32), ouv.view(np.int32))
        np.testing.assert_array_equal(pdf.cpu().numpy().view(np.int32), opdf.view(np.int32))


@pytest.mark.parametrize("uniform", [False, True])
def test_adaptive_environment_light(ctx, uniform):
    """illuminateEnvironmentLight1 / getRadianceEnvironmentLight1 /
    accumLightGrid against the oracle on a non-zero camera position: grid
    entries, directions, texels, pdfs and luminances bit-exact; the float
    atomic accumulation exact on a constant map (equal addends sum the same
    in any order), within 1e-5 of a float64 sum otherwise."""
    print(f"[case] adaptive env light uniform={uniform}", flush=True)
    w, h = 1024, 512
    rng = np.random.default_rng(31)
    rgba = np.ones((h, w, 4), np.float32) if uniform else _env_map(w, h)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    grid = api.CAdaptiveEnvironment(ctx)
    accum = (rng.random((1600, 1600)) ** 3 + 0.01).astype(np.float32)
    grid.write_accum(accum)
    grid.update()
    cam = np.array([0.5, 1.0, -4.0], np.float32)
    grid.setCameraPosition(cam)
    n = 200000
    hit = np.concatenate([rng.uniform(-5, 5, (n, 3)), np.zeros((n, 1))], 1).astype(np.float32)
    u = rng.random((n, 4)).astype(np.float32)
    d, c, pw, lo, gi = env.illuminate_adaptive(grid, ctx.to_device(hit), ctx.to_device(u))
    lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), w, h)
    _, _, gsat = grid.read_arrays()
    od, oc, opw, olo, ogi, _ = orc.env_illuminate_adaptive(rgba.reshape(-1, 4), lum, sat, w, h, accum.reshape(-1),
                                                           gsat.reshape(-1), cam, hit, u)
    gi, d, c, pw, lo = gi.cpu().numpy(), d.cpu().numpy(), c.cpu().numpy(), pw.cpu().numpy(), lo.cpu().numpy()
    assert bits("grid index", gi, ogi) + bits("dir", d, od) + bits("contrib", c, oc) + bits("pdf", pw, opw) \
        + bits("lum", lo, olo) == 0

    # escaped rays: radiance with accumulation into a cleared accum
    dirs = rng.normal(size=(n, 3))
    dirs = np.concatenate([dirs / np.linalg.norm(dirs, axis=1, keepdims=True), np.zeros((n, 1))], 1).astype(np.float32)
    grid.clearBuffers()
    c2, pa, lo2, gi2 = env.radiance_adaptive(grid, ctx.to_device(dirs), ctx.to_device(hit))
    acc_gpu = grid.read_arrays()[0].reshape(-1)
    oacc = np.zeros(1600 * 1600, np.float32)
    oc2, opa, olo2, ogi2 = orc.env_radiance_adaptive(rgba.reshape(-1, 4), lum, sat, w, h, cam, dirs, hit, accum=oacc)
    gi2, pa, lo2 = gi2.cpu().numpy(), pa.cpu().numpy(), lo2.cpu().numpy()
    assert bits("grid index", gi2, ogi2) + bits("pdf", pa, opa) + bits("lum", lo2, olo2) \
        + bits("contrib", c2.cpu().numpy(), oc2) == 0
    same = np.ones(n, bool)
    # the accum holds exactly the GPU's own (entry, luminance) pairs
    want = np.zeros(1600 * 1600, np.float32 if uniform else np.float64)
    np.add.at(want, gi2, lo2.astype(want.dtype))
    if uniform:
        np.testing.assert_array_equal(acc_gpu.view(np.int32), want.view(np.int32))
        # and the oracle's accum wherever no mismatched sample landed
        clean = np.ones(1600 * 1600, bool)
        clean[gi2[~same]] = False
        clean[ogi2[~same]] = False
        np.testing.assert_array_equal(acc_gpu[clean].view(np.int32), oacc[clean].view(np.int32))
    else:
        np.testing.assert_allclose(acc_gpu, want, rtol=1e-5, atol=0)
    assert (acc_gpu > 0).sum() > 1000
