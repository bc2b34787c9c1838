"""Parity at the full workload sizes (config 1: 1920x1080 x 16 spp over
1,048,578 triangles; config 2: 64M rays through 4M triangles).  The oracle
cannot trace whole jobs in seconds, so these tests run the full job on the
GPU and check (a) a seeded random sample of it against the oracle at the
usual bars, and (b) size-independent properties of the whole output:
counts, compaction order, agreement between traversal modes.

Bars: bit-exact everywhere: config-1 pixels (both lighting paths, the
light-grid guided one across frames with grid updates, its grid sums too);
config-2 primitive ids, t and occlusion."""
import os

import numpy as np
import pytest
import torch

from mambatracer_amd import _lib, api
from mambatracer_amd import workloads as W
from oracle import orc_py as orc

pytestmark = pytest.mark.gpu


def test_c1_full_render_sampled_pixels(ctx):
    print("[case] c1 full render, sampled pixels", flush=True)
    points, normals, faces, nodes, bounds, depth = W.c1_scene()
    mats = W.c1_materials()
    rgba = W.env_map()
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    job = W.RenderJob()
    img, st = api.render(ctx, sc, ctx.to_device(mats), env, job)
    img = img.cpu().numpy()[:, :3]
    w, h, spp = job.width, job.height, job.spp
    assert st.bounce_counts[0] == w * h * spp and int(st.paths) == w * h * spp
    assert np.isfinite(img).all() and img.mean() > 0
    # 300 seeded random pixels, all of their samples through the oracle
    rng = np.random.default_rng(11)
    px = np.sort(rng.choice(w * h, 300, replace=False))
    first = orc.pixel_first_paths(w, h, spp)[px]
    pids = (first[:, None] + np.arange(spp)[None, :]).reshape(-1).astype(np.int32)
    lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), 1024, 512)
    L, _ = orc.render_paths(points, normals, faces, nodes, bounds, mats, (rgba.reshape(-1, 4), lum, sat, 1024, 512),
                            job, np.sort(pids))
    acc = np.zeros((len(px), 3), np.float32)
    for s in range(spp):
        acc = (acc + L[first + s, :3]).astype(np.float32)
    oimg = (acc * np.float32(1.0 / spp)).astype(np.float32)
    got = img[px]
    bad = np.any(got.view(np.int32) != oimg.view(np.int32), axis=1)
    print(f"pixels differing from the oracle: {int(bad.sum())} of {len(px)} {px[bad][:6].tolist()}")
    assert not bad.any()


def _threads():
    return max(1, min(32, os.cpu_count() or 1))


@pytest.mark.parametrize("tile_pdf", [False, True])
def test_c1_full_guided_frames(ctx, tile_pdf):
    """Config 1 with the environment sampled through the light grid:
    frames 0, 1, 2 of one view with the grid updated after frames 0 and 1
    (CDeviceGI.java:263-264).  The oracle renders frames 0 and 1 whole (their
    grid sums feed the updates) and 300 seeded pixels of frame 2; the grid's
    accum after each frame and frame 2's pixels must be bit-exact."""
    print(f"[case] c1 full guided frames tile_pdf={tile_pdf}", flush=True)
    points, normals, faces, nodes, bounds, depth = W.c1_scene()
    mats = W.c1_materials()
    rgba = W.env_map()
    lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), 1024, 512)
    oenv = (rgba.reshape(-1, 4), lum, sat, 1024, 512)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    flags = _lib.MT_RENDER_LIGHTGRID_ENV | (_lib.MT_RENDER_LIGHTGRID_TILE_PDF if tile_pdf else 0)
    job = W.RenderJob(flags=flags)
    tile = api.lightgrid_seed_tile(lum, 1024, 512)
    grid = api.CAdaptiveEnvironment(ctx)
    grid.setHDRLuminance(tile)
    grid.setCameraPosition(job.pos)
    env.attachLightGrid(grid)
    og = orc.LightGrid(tile, cam=job.pos)
    dm = ctx.to_device(mats)
    w, h, spp = job.width, job.height, job.spp
    for frame in range(3):
        job.seed = 2 + frame
        img, st = api.render(ctx, sc, dm, env, job)
        img = img.cpu().numpy()[:, :3]
        assert np.isfinite(img).all() and img.mean() > 0
        if frame < 2:
            L, rays = orc.render_paths(points, normals, faces, nodes, bounds, mats, oenv, job,
                                       np.arange(w * h * spp, dtype=np.int32), nthreads=_threads(), grid=og,
                                       grid_mode=2 if tile_pdf else 1)
            og.fold()
            oimg = orc.resolve(L, w, h, spp)
            del L
            acc = grid.read_arrays()[0].reshape(-1)
            nbad_acc = int((acc.view(np.int32) != og.accum.view(np.int32)).sum())
            bad = np.any(img.view(np.int32) != oimg.view(np.int32), axis=1)
            print(f"frame {frame}: rays gpu {st.extension_rays}+{st.shadow_rays} oracle {rays[0]}+{rays[1]}; "
                  f"pixels differing {int(bad.sum())}; accum entries differing {nbad_acc}", flush=True)
            assert (int(st.extension_rays), int(st.shadow_rays)) == (int(rays[0]), int(rays[1]))
            assert not bad.any() and nbad_acc == 0
            grid.update()
            og.update()
        else:
            rng = np.random.default_rng(12)
            px = np.sort(rng.choice(w * h, 300, replace=False))
            first = orc.pixel_first_paths(w, h, spp)[px]
            pids = (first[:, None] + np.arange(spp)[None, :]).reshape(-1).astype(np.int32)
            L, _ = orc.render_paths(points, normals, faces, nodes, bounds, mats, oenv, job, np.sort(pids),
                                    nthreads=_threads(), grid=og, grid_mode=2 if tile_pdf else 1)
            acc3 = np.zeros((len(px), 3), np.float32)
            for s in range(spp):
                acc3 = (acc3 + L[first + s, :3]).astype(np.float32)
            oimg = (acc3 * np.float32(1.0 / spp)).astype(np.float32)
            bad = np.any(img[px].view(np.int32) != oimg.view(np.int32), axis=1)
            print(f"frame 2: sampled pixels differing {int(bad.sum())} of {len(px)}", flush=True)
            assert not bad.any()


def test_c1_full_tail_matches_wavefront(ctx):
    """Config 1 (guided light, 1920x1080 x 16 spp) with bounces 3.. in one
    k_tail launch (MT_RENDER_TAIL_FROM(3), the benchmark's setting) against
    the per-bounce wavefront, from the same grid state: image, ray counts,
    per-bounce path counts and the grid's sums are bit-identical."""
    print("[case] c1 full tail vs wavefront", flush=True)
    points, normals, faces, nodes, bounds, depth = W.c1_scene()
    rgba = W.env_map()
    lum, _ = orc.env_lum_sat(rgba.reshape(-1, 4), 1024, 512)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    tile = api.lightgrid_seed_tile(lum, 1024, 512)
    grid = api.CAdaptiveEnvironment(ctx)
    job = W.RenderJob(flags=_lib.MT_RENDER_LIGHTGRID_ENV, seed=7)
    grid.setCameraPosition(job.pos)
    env.attachLightGrid(grid)
    dm = ctx.to_device(W.c1_materials())
    out = []
    for flags in (_lib.MT_RENDER_LIGHTGRID_ENV, _lib.MT_RENDER_LIGHTGRID_ENV | _lib.MT_RENDER_TAIL_FROM(3)):
        grid.clearBuffers()
        grid.setHDRLuminance(tile)
        job.flags = flags
        img, st = api.render(ctx, sc, dm, env, job)
        torch.cuda.synchronize()
        out.append((img.cpu().numpy(), st, grid.read_arrays()[0].copy()))
    (img0, st0, acc0), (img1, st1, acc1) = out
    bad = np.any(img0.view(np.int32) != img1.view(np.int32), axis=1)
    print(f"rays {st0.extension_rays}+{st0.shadow_rays} vs {st1.extension_rays}+{st1.shadow_rays}; "
          f"per bounce {list(st0.bounce_counts)[:8]}; pixels differing {int(bad.sum())}", flush=True)
    assert not bad.any()
    assert (int(st0.extension_rays), int(st0.shadow_rays)) == (int(st1.extension_rays), int(st1.shadow_rays))
    assert list(st0.bounce_counts) == list(st1.bounce_counts)
    assert np.array_equal(acc0.view(np.int32), acc1.view(np.int32)) and acc0.sum() > 0


def test_c2_full_batch_sampled_rays_and_whole_batch_properties(ctx):
    print("[case] c2 full batch (64M rays), sampled rays", flush=True)
    points, faces, nodes, bounds, depth = W.c2_scene()
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth)
    n = 64 << 20
    o, d = api.gen_random_rays(3, 0, n, bounds[0, :3], bounds[0, 4:7], nthreads=16)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    prim, t, _ = api.intersect_primitives(ctx, sc, do, dd)                               # one ray per lane
    lane_maps = (_lib.MT_TRACE_PACKED_REFILL, _lib.MT_TRACE_PACKED_POOL, _lib.MT_TRACE_PACKED_BINNED)
    for m in lane_maps:  # every lane mapping gives every ray's result of one ray per lane
        prim_r, t_r, _ = api.intersect_primitives(ctx, sc, do, dd, mode=m)
        assert torch.equal(prim, prim_r) and torch.equal(t.view(torch.int32), t_r.view(torch.int32)), m
        del prim_r, t_r
    light = W.c2_light(bounds)
    so, sd, valid = api.shadow_rays(ctx, do, dd, prim, t, light)
    occ, _ = api.intersect_occlusion(ctx, sc, so, sd)
    for m in lane_maps:
        occ_r, _ = api.intersect_occlusion(ctx, sc, so, sd, mode=m)
        assert torch.equal(occ, occ_r), m
        del occ_r
    # stable compaction of the hit rays: the index list is exactly the
    # sorted positions of prim >= 0 and the payload follows it
    idx, cnt, (co,) = api.compact(ctx, prim, flag_mode=1, payloads=[dd])
    hits = torch.nonzero(prim >= 0).squeeze(1).to(torch.int32)
    k = int(cnt.item())
    assert k == hits.numel() and torch.equal(idx[:k], hits)
    assert torch.equal(co[:k], dd[hits.long()])
    ctx.check_scans()
    # 40,000 seeded random rays against the oracle
    rng = np.random.default_rng(12)
    sample = np.sort(rng.choice(n, 40_000, replace=False))
    oprim, ot, _, _ = orc.trace_closest(points, faces, nodes, bounds, o[sample], d[sample])
    np.testing.assert_array_equal(prim.cpu().numpy()[sample], oprim)
    np.testing.assert_array_equal(t.cpu().numpy()[sample].view(np.int32), ot.view(np.int32))
    s_o, s_d = so.cpu().numpy()[sample], sd.cpu().numpy()[sample]
    oocc, _, _ = orc.trace_any(points, faces, nodes, bounds, s_o, s_d)
    np.testing.assert_array_equal(occ.cpu().numpy()[sample], oocc)
    print(f"hits {k} of {n}; sample: {int((oprim >= 0).sum())} hits, {int(oocc.sum())} occluded")
