"""Captured render (mt_render_graph): the bounce loop recorded once into a
HIP graph and replayed per frame.  Bar: every replay's image and ray counts
equal mt_render's for the same seed bit for bit (the graph replays the same
launches), with eager library work and synchronisations between replays, and
one replay against the oracle's restatement at the render tolerance."""
import numpy as np
import pytest
import torch

from mambatracer_amd import _lib, api
from mambatracer_amd import workloads as W
from oracle import orc_py as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene_small(ctx):
    points, normals, faces, nodes, bounds, depth = W.c1_scene(count=32)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    mats = W.c1_materials()
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    return sc, mats, ctx.to_device(mats), env, (points, normals, faces, nodes, bounds)


def _job(seed=1, flags=0, **kw):
    args = dict(width=128, height=72, spp=4, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0))
    args.update(kw)
    return W.RenderJob(seed=seed, flags=flags, **args)


def _eager(ctx, sc, dm, env, job, **kw):
    img, st = api.render(ctx, sc, dm, env, job, **kw)
    return img.cpu().numpy(), int(st.extension_rays), int(st.shadow_rays), list(st.bounce_counts)


def test_graph_replays_equal_eager_renders(ctx, scene_small):
    print("[case] render graph replays", flush=True)
    sc, _, dm, env, _ = scene_small
    g = api.RenderGraph(ctx, sc, dm, env, _job(seed=1))
    try:
        for seed in (11, 12, 11, 7):
            g.launch(seed)
            st = g.stats()   # synchronises
            got = g.image.cpu().numpy()
            # eager render of the same seed, between replays
            ref, ext, sh, bc = _eager(ctx, sc, dm, env, _job(seed=seed))
            torch.cuda.synchronize()
            np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
            assert int(st.extension_rays) == ext and int(st.shadow_rays) == sh
            assert list(st.bounce_counts) == bc
        # back-to-back replays without a sync between them: the last one wins
        for seed in (3, 4, 5):
            g.launch(seed)
        torch.cuda.synchronize()
        ref, *_ = _eager(ctx, sc, dm, env, _job(seed=5))
        np.testing.assert_array_equal(g.image.cpu().numpy().view(np.int32), ref.view(np.int32))
    finally:
        g.close()


def test_graph_matches_oracle(ctx, scene_small):
    print("[case] render graph vs oracle", flush=True)
    sc, mats, dm, env, (points, normals, faces, nodes, bounds) = scene_small
    job = _job(seed=21, width=96, height=54)
    g = api.RenderGraph(ctx, sc, dm, env, job)
    try:
        g.launch(21)
        st = g.stats()
        img = g.image.cpu().numpy()[:, :3]
    finally:
        g.close()
    rgba = W.env_map()
    lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), 1024, 512)
    npix = job.width * job.height
    pids = np.arange(npix * job.spp, dtype=np.int32)
    L, rays = orc.render_paths(points, normals, faces, nodes, bounds, mats, (rgba.reshape(-1, 4), lum, sat, 1024, 512),
                               job, pids)
    oimg = orc.resolve(L, job.width, job.height, job.spp)
    ext, sh = int(st.extension_rays), int(st.shadow_rays)
    assert abs(ext - rays[0]) <= 0.005 * rays[0] and abs(sh - rays[1]) <= 0.005 * rays[1] + 2
    rmse = np.sqrt(np.mean((img - oimg) ** 2)) / np.sqrt(np.mean(oimg ** 2))
    print(f"relative RMSE {rmse:.2e}")
    assert rmse <= 1e-3


def test_graph_survives_context_workspace_regrowth(ctx, scene_small):
    """The graph owns its look-back workspace: a later, larger compaction on
    the same context regrows (and frees) the context's workspace, and replays
    still produce the eager image."""
    print("[case] render graph after workspace regrowth", flush=True)
    sc, _, dm, env, _ = scene_small
    g = api.RenderGraph(ctx, sc, dm, env, _job(seed=2))
    try:
        g.launch(2)
        torch.cuda.synchronize()
        first = g.image.cpu().numpy().copy()
        # A compaction declared with capacity 2^30 (one status word per
        # 8192-element tile: just over the 1 MiB minimum workspace) regrows
        # the context's workspace; its device length is 5000, so only the
        # first tile reads flags.
        n = 5000
        flags = torch.ones(n, dtype=torch.int32, device=ctx.torch_device)
        length = ctx.scalar_i32(n)
        cnt = ctx.scalar_i32(0)
        api.call("mt_compact", ctx.handle, flags.data_ptr(), 0, length.data_ptr(), 1 << 30, None,
                 cnt.data_ptr(), None, 0)
        assert int(cnt.item()) == n
        g.launch(2)
        g.stats()
        np.testing.assert_array_equal(g.image.cpu().numpy().view(np.int32), first.view(np.int32))
    finally:
        g.close()


def test_graph_mesh_lights_and_traversal_stats(ctx, scene_small):
    print("[case] render graph mesh lights + traversal stats", flush=True)
    sc, mats, _, env, _ = scene_small
    m = mats.copy()
    m[:4, 28:31] = [[1.0, 0.9, 0.7]]
    m[:4, 32] = 1.0
    m[:4, 33] = 5.0
    dm = ctx.to_device(m)
    lf, nl = api.mesh_lights(ctx, sc, dm)
    job = _job(seed=9, flags=_lib.MT_RENDER_PICKED_LIGHT | _lib.MT_RENDER_TRAVERSAL_STATS)
    g = api.RenderGraph(ctx, sc, dm, env, job, light_faces=lf, num_lights=nl)
    try:
        g.launch(9)
        st = g.stats()
        got = g.image.cpu().numpy()
    finally:
        g.close()
    img, est = api.render(ctx, sc, dm, env, job, light_faces=lf, num_lights=nl)
    np.testing.assert_array_equal(got.view(np.int32), img.cpu().numpy().view(np.int32))
    assert list(st.closest_ilp) == list(est.closest_ilp) and list(st.any_ilp) == list(est.any_ilp)
    assert st.closest_ilp[0] > 0


def test_graph_rejects_event_profiling(ctx, scene_small):
    sc, _, dm, env, _ = scene_small
    with pytest.raises(RuntimeError, match="PROFILE"):
        api.RenderGraph(ctx, sc, dm, env, _job(flags=_lib.MT_RENDER_PROFILE))
