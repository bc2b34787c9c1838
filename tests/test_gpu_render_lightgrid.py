"""The bounce loop with the light-grid guided environment light
(MT_RENDER_LIGHTGRID_ENV: illuminateEnvironmentLight1 for the light sample,
getRadianceEnvironmentLight1 + accumLightGrid for escaped rays and unoccluded
light samples, Light.cl:35-38,140-268, Render.cl:458-459) against the
oracle's loop, over a sequence of frames with the grid updated between them
as CDeviceGI.loop does for frames 0-4 (CDeviceGI.java:263-264,
CAdaptiveEnvironment.update :135-145).

Both sides use the same transcendental functions and fixed-point grid sums,
so the bar is bit-exact: images, ray counts, the grid's accum after each
frame, and func / SAT after each update.  Both pdf variants: the reference's
in-tile pdf, and the tile pick's pdf applied."""
import numpy as np
import pytest
import torch

from mambatracer_amd import _lib, api
from mambatracer_amd import workloads as W
from oracle import orc_py as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_small():
    return W.c1_scene(count=32)


def _setup(ctx, c1_small, job):
    points, normals, faces, nodes, bounds, depth = c1_small
    mats = W.c1_materials()
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    rgba = W.env_map()
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), 1024, 512)
    oenv = (rgba.reshape(-1, 4), lum, sat, 1024, 512)
    tile = api.lightgrid_seed_tile(lum, 1024, 512)
    grid = api.CAdaptiveEnvironment(ctx)
    grid.setHDRLuminance(tile)
    grid.setCameraPosition(job.pos)
    env.attachLightGrid(grid)
    og = orc.LightGrid(tile, cam=job.pos)
    return sc, mats, env, oenv, grid, og


def _diff(name, a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    bad = a.view(np.int32) != b.view(np.int32)
    n = int(bad.sum())
    if n:
        idx = np.argwhere(bad)[:8]
        print(f"{name}: {n} of {bad.size} differ, first {idx.tolist()}: "
              f"{a[tuple(idx.T)].tolist()} vs {b[tuple(idx.T)].tolist()}", flush=True)
    return n


@pytest.mark.parametrize("tile_pdf,tail", [(False, 0), (True, 0), (False, 2), (True, 1)])
def test_guided_frames_match_oracle(ctx, c1_small, tile_pdf, tail):
    """tail > 0: bounces tail.. run in one k_tail launch (MT_RENDER_TAIL_FROM)."""
    flags = _lib.MT_RENDER_LIGHTGRID_ENV | (_lib.MT_RENDER_LIGHTGRID_TILE_PDF if tile_pdf else 0)
    flags |= _lib.MT_RENDER_TAIL_FROM(tail)
    job = W.RenderJob(width=96, height=54, spp=4, max_bounces=6, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0),
                      flags=flags)
    print(f"[case] guided frames tile_pdf={tile_pdf} tail={tail}", flush=True)
    points, normals, faces, nodes, bounds, depth = c1_small
    sc, mats, env, oenv, grid, og = _setup(ctx, c1_small, job)
    dm = ctx.to_device(mats)
    pids = np.arange(job.width * job.height * job.spp, dtype=np.int32)
    for frame in range(3):
        job.seed = 40 + frame
        img, st = api.render(ctx, sc, dm, env, job)
        torch.cuda.synchronize()
        L, rays = orc.render_paths(points, normals, faces, nodes, bounds, mats, oenv, job, pids, grid=og,
                                   grid_mode=2 if tile_pdf else 1)
        og.fold()
        oimg = orc.resolve(L, job.width, job.height, job.spp)
        accum = grid.read_arrays()[0].reshape(-1)
        print(f"frame {frame}: rays gpu {st.extension_rays}+{st.shadow_rays} oracle {rays[0]}+{rays[1]}, "
              f"accum total {accum.astype(np.float64).sum():.6g}", flush=True)
        assert (int(st.extension_rays), int(st.shadow_rays)) == (int(rays[0]), int(rays[1]))
        assert _diff("image", img.cpu().numpy()[:, :3], oimg) == 0
        assert _diff("accum", accum, og.accum) == 0
        assert og.accum.sum() > 0
        grid.update()
        og.update()
        _, func, gsat = grid.read_arrays()
        assert _diff("func", func.reshape(-1), og.func) == 0
        assert _diff("sat", gsat.reshape(-1), og.sat) == 0
        grid.buildSAT()  # reading func/sat hands the arrays out; rebuild the sampling tables


def test_guided_flags_need_a_grid(ctx, c1_small):
    points, normals, faces, nodes, bounds, depth = c1_small
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    dm = ctx.to_device(W.c1_materials())
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    job = W.RenderJob(width=8, height=8, spp=1, flags=_lib.MT_RENDER_LIGHTGRID_ENV)
    with pytest.raises(_lib.MambartError, match="light grid"):
        api.render(ctx, sc, dm, env, job)
    job.flags = _lib.MT_RENDER_LIGHTGRID_TILE_PDF
    with pytest.raises(_lib.MambartError, match="TILE_PDF"):
        api.render(ctx, sc, dm, env, job)


@pytest.mark.parametrize("tail", [0, 3])
def test_guided_render_graph_replays_match_render(ctx, c1_small, tail):
    """A captured guided render replays the same image and grid sums as
    mt_render from the same grid state (eager work between replays); with
    tail > 0 the capture holds the one-launch tail (MT_RENDER_TAIL_FROM)."""
    print(f"[case] guided render graph tail={tail}", flush=True)
    job = W.RenderJob(width=64, height=40, spp=2, max_bounces=5, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0),
                      flags=_lib.MT_RENDER_LIGHTGRID_ENV | _lib.MT_RENDER_LIGHTGRID_TILE_PDF
                      | _lib.MT_RENDER_TAIL_FROM(tail))
    sc, mats, env, oenv, grid, og = _setup(ctx, c1_small, job)
    dm = ctx.to_device(mats)
    g = api.RenderGraph(ctx, sc, dm, env, job)
    for rep in range(3):
        grid.clearBuffers()
        grid.setHDRLuminance(api.lightgrid_seed_tile(oenv[1], 1024, 512))
        img, _ = api.render(ctx, sc, dm, env, job)
        a_eager = grid.read_arrays()[0].copy()
        grid.setHDRLuminance(api.lightgrid_seed_tile(oenv[1], 1024, 512))
        torch.zeros(1000, device="cuda").sum().item()  # eager work between replays
        g.launch(job.seed)
        torch.cuda.synchronize()
        a_graph = grid.read_arrays()[0]
        assert _diff("graph image", g.image.cpu().numpy(), img.cpu().numpy()) == 0
        assert _diff("graph accum", a_graph, a_eager) == 0
    g.close()
