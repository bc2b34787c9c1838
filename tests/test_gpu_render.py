"""The wavefront bounce loop on the GPU against the oracle's one-path-at-a-
time restatement of the same loop (config-1 scene, reduced image).

Bar: bit-exact images and equal ray counts.  Both sides run the same IEEE
operations in the same order, the transcendental functions included
(mt_xmath.h), and every result is indexed by path id, so nothing depends on
scheduling.  Mismatching pixels, if any, are listed."""
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


def image_mismatches(img, oimg):
    bad = np.any(img.view(np.int32) != oimg.view(np.int32), axis=1)
    if bad.any():
        idx = np.flatnonzero(bad)[:6]
        print(f"{int(bad.sum())} pixels differ, e.g. {idx.tolist()}: {img[idx].tolist()} vs {oimg[idx].tolist()}",
              flush=True)
    return int(bad.sum())


def _render_both(ctx, c1_small, job, with_env=True):
    points, normals, faces, nodes, bounds, depth = c1_small
    mats = W.c1_materials()
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    rgba = W.env_map()
    env = None
    oenv = None
    if with_env:
        env = api.CEnvironment(ctx)
        env.setEnvironmentMap(rgba)
        lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), 1024, 512)
        oenv = (rgba.reshape(-1, 4), lum, sat, 1024, 512)
    img, st = api.render(ctx, sc, ctx.to_device(mats), env, job)
    npix = job.width * job.height
    pids = np.arange(npix * job.spp, dtype=np.int32)
    L, rays = orc.render_paths(points, normals, faces, nodes, bounds, mats, oenv, job, pids)
    oimg = orc.resolve(L, job.width, job.height, job.spp)
    return img.cpu().numpy()[:, :3], st, oimg, rays


def test_render_matches_oracle(ctx, c1_small):
    print("[case] render c1 small", flush=True)
    job = W.RenderJob(width=160, height=90, spp=4, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0))
    img, st, oimg, rays = _render_both(ctx, c1_small, job)
    ext, sh = int(st.extension_rays), int(st.shadow_rays)
    print(f"gpu rays {ext}+{sh}, oracle {rays[0]}+{rays[1]}")
    assert (ext, sh) == (int(rays[0]), int(rays[1]))
    assert image_mismatches(img, oimg) == 0
    assert np.isfinite(img).all() and img.mean() > 0


@pytest.mark.parametrize("flags", [0, _lib.MT_RENDER_PICKED_LIGHT])
def test_render_mesh_lights_matches_oracle(ctx, c1_small, flags):
    """Emitter hits (LightHitPass emitter branch) and DirectLight on area
    lights picked from the mesh-light list, beside the environment."""
    print(f"[case] render mesh lights flags={flags}", flush=True)
    points, normals, faces, nodes, bounds, depth = c1_small
    mats = W.c1_materials()
    mats[:4, 28:31] = [[1.0, 0.9, 0.7]]   # spheres 0-3 emit
    mats[:4, 32] = 1.0
    mats[:4, 33] = 5.0
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    dm = ctx.to_device(mats)
    lf, nl = api.mesh_lights(ctx, sc, dm)
    rgba = W.env_map()
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    job = W.RenderJob(width=96, height=54, spp=4, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0), flags=flags)
    img, st = api.render(ctx, sc, dm, env, job, light_faces=lf, num_lights=nl)
    torch.cuda.synchronize()
    n = int(nl.item())
    assert n == 4 * 4096
    hlf = lf.cpu().numpy()[:n].copy()
    lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), 1024, 512)
    npix = job.width * job.height
    L, rays = orc.render_paths(points, normals, faces, nodes, bounds, mats, (rgba.reshape(-1, 4), lum, sat, 1024, 512),
                               job, np.arange(npix * job.spp, dtype=np.int32), light_faces=hlf)
    oimg = orc.resolve(L, job.width, job.height, job.spp)
    img = img.cpu().numpy()[:, :3]
    ext, sh = int(st.extension_rays), int(st.shadow_rays)
    assert (ext, sh) == (int(rays[0]), int(rays[1]))
    assert image_mismatches(img, oimg) == 0


def test_render_is_deterministic_and_independent_of_batch(ctx, c1_small):
    """Same seed -> bit-identical image; each sample's paths do not depend on the
    other samples in the wavefront (spp=1 render == sample 0 of an spp=3 render)."""
    points, normals, faces, nodes, bounds, depth = c1_small
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    dm = ctx.to_device(W.c1_materials())
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    job = W.RenderJob(width=128, height=72, spp=3, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0))
    a, _ = api.render(ctx, sc, dm, env, job)
    b, _ = api.render(ctx, sc, dm, env, job)
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_render_without_environment_and_reference_layout(ctx, c1_small):
    job = W.RenderJob(width=64, height=36, spp=2, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0), mode=1)
    img, st, oimg, rays = _render_both(ctx, c1_small, job, with_env=False)
    assert (img == 0).all() and (oimg == 0).all()   # nothing emits: black image
    assert int(st.shadow_rays) == 0 and rays[1] == 0


def test_render_stats_counters(ctx, c1_small):
    points, normals, faces, nodes, bounds, depth = c1_small
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    dm = ctx.to_device(W.c1_materials())
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    job = W.RenderJob(width=64, height=64, spp=1, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0),
                      flags=_lib.MT_RENDER_TRAVERSAL_STATS | _lib.MT_RENDER_PROFILE)
    _, st = api.render(ctx, sc, dm, env, job)
    assert st.bounce_counts[0] == 64 * 64
    assert st.closest_ilp[0] > 0 and st.closest_launches == job.max_bounces
    assert st.closest_ms > 0


@pytest.mark.parametrize("mesh_lights", [False, True])
def test_fused_and_staged_loops_agree(ctx, c1_small, mesh_lights):
    """The default loop fuses the camera pass into bounce 0's traversal and
    the direct-light add into the any-hit kernel; with traversal counters on,
    or another traversal layout, every stage is a separate launch.  Same
    paths, so images, ray counts and per-bounce counts are bit-identical."""
    print(f"[case] fused vs staged loop, mesh lights {mesh_lights}", flush=True)
    points, normals, faces, nodes, bounds, depth = c1_small
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    mats = W.c1_materials()
    lf = nl = None
    if mesh_lights:
        mats[:4, 28:31] = [[1.0, 0.9, 0.7]]
        mats[:4, 32] = 1.0
        mats[:4, 33] = 5.0
    dm = ctx.to_device(mats)
    if mesh_lights:
        lf, nl = api.mesh_lights(ctx, sc, dm)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    base = dict(width=100, height=60, spp=3, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0), seed=5)
    runs = [W.RenderJob(**base),                                           # fused
            W.RenderJob(**base, mode=_lib.MT_TRACE_PACKED_SORTED),         # fused, direction-sorted lanes
            W.RenderJob(**base, mode=_lib.MT_TRACE_PACKED_POOL),           # fused, wave-pool lanes
            W.RenderJob(**base, flags=_lib.MT_RENDER_TAIL_FROM(1)),        # fused, bounces 1.. in one launch
            W.RenderJob(**base, flags=_lib.MT_RENDER_TAIL_FROM(3)),        # fused, bounces 3.. in one launch
            W.RenderJob(**base, flags=_lib.MT_RENDER_TAIL_FROM(W.RenderJob().max_bounces - 1)),  # last bounce only
            W.RenderJob(**base, flags=_lib.MT_RENDER_TAIL_FROM(W.RenderJob().max_bounces)),      # past the end: off
            W.RenderJob(**base, mode=_lib.MT_TRACE_PACKED_POOL,            # wave pools, then the tail
                        flags=_lib.MT_RENDER_TAIL_FROM(2)),
            W.RenderJob(**base, flags=_lib.MT_RENDER_TRAVERSAL_STATS),     # staged, same traversal
            W.RenderJob(**base, mode=_lib.MT_TRACE_PACKED_NO_STACK),       # staged, parent-link pops
            W.RenderJob(**base, mode=_lib.MT_TRACE_REFERENCE)]             # staged, reference arrays
    out = [api.render(ctx, sc, dm, env, j, light_faces=lf, num_lights=nl) for j in runs]
    ref_img, ref_st = out[0]
    for img, st in out[1:]:
        np.testing.assert_array_equal(img.cpu().numpy().view(np.int32), ref_img.cpu().numpy().view(np.int32))
        assert int(st.extension_rays) == int(ref_st.extension_rays)
        assert int(st.shadow_rays) == int(ref_st.shadow_rays)
        assert list(st.bounce_counts) == list(ref_st.bounce_counts)


@pytest.mark.parametrize("mesh_lights", [False, True])
def test_tail_with_many_live_paths(ctx, mesh_lights):
    """A dense triangle soup in front of the camera keeps many paths alive
    into the late bounces (a fifth at bounce 3, against 0.2% in the open
    config-1 scene), so the one-launch tail carries thousands of paths per
    segment sweep and refills every lane many times.  Tail from bounce 1, 3
    and 6 against the per-bounce launches: same image bits, ray counts and
    per-bounce counts."""
    print(f"[case] tail on a soup, mesh lights {mesh_lights}", flush=True)
    points, faces, nodes, bounds, depth = W.soup_scene(200_000, 61)
    mats = W.c1_materials()
    lf = nl = None
    if mesh_lights:
        emit = W.material_records(np.full((1, 3), 0.5), 0.04, 0.5, 0.5, emission=np.full((1, 3), 1.0),
                                  emission_power=20.0)
        mats = np.concatenate([mats, emit])
        faces[::400, 9] = mats.shape[0] - 1
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth)
    dm = ctx.to_device(mats)
    if mesh_lights:
        lf, nl = api.mesh_lights(ctx, sc, dm)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    job = W.soup_job(W.JobSpec(0, 0, 128, 4, 6100 + int(mesh_lights)))
    out = []
    for tail in (0, 1, 3, job.max_bounces - 2):
        job.flags = _lib.MT_RENDER_TAIL_FROM(tail)
        img, st = api.render(ctx, sc, dm, env, job, light_faces=lf, num_lights=nl)
        out.append((img.cpu().numpy(), st))
    ref_img, ref_st = out[0]
    print(f"paths per bounce {list(ref_st.bounce_counts)}", flush=True)
    for img, st in out[1:]:
        np.testing.assert_array_equal(img.view(np.int32), ref_img.view(np.int32))
        assert int(st.extension_rays) == int(ref_st.extension_rays)
        assert int(st.shadow_rays) == int(ref_st.shadow_rays)
        assert list(st.bounce_counts) == list(ref_st.bounce_counts)
    assert np.isfinite(ref_img).all()
    paths = 128 * 128 * 4
    assert ref_st.bounce_counts[1] > paths // 2 and ref_st.bounce_counts[3] > paths // 8


@pytest.mark.parametrize("w,h,spp", [(1, 1, 1), (7, 3, 5), (17, 9, 1)])
def test_tiny_jobs_fused_and_staged_agree(ctx, c1_small, w, h, spp):
    """Jobs far smaller than one work-group or one queue segment: the fused
    loop (appended, segmented path queues) and the staged loop (compaction)
    give the same image bits and ray counts, and the counts add up."""
    print(f"[case] tiny job {w}x{h}x{spp}", flush=True)
    points, normals, faces, nodes, bounds, depth = c1_small
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    dm = ctx.to_device(W.c1_materials())
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    base = dict(width=w, height=h, spp=spp, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0), seed=9)
    (img_f, st_f), (img_s, st_s), (img_t, st_t) = [
        api.render(ctx, sc, dm, env, W.RenderJob(**base, **extra))
        for extra in ({}, {"flags": _lib.MT_RENDER_TRAVERSAL_STATS}, {"flags": _lib.MT_RENDER_TAIL_FROM(1)})]
    for img, st in ((img_s, st_s), (img_t, st_t)):
        np.testing.assert_array_equal(img_f.cpu().numpy().view(np.int32), img.cpu().numpy().view(np.int32))
        assert list(st_f.bounce_counts) == list(st.bounce_counts)
        assert int(st_f.shadow_rays) == int(st.shadow_rays)
    assert st_f.bounce_counts[0] == w * h * spp == int(st_f.paths)
    assert int(st_f.extension_rays) == sum(st_f.bounce_counts)
    assert np.isfinite(img_f.cpu().numpy()).all()
