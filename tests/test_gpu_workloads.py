"""The multi-job workloads at their full sizes.

Config 3: the whole 256-job generator (image sides log-uniform 64-1024,
1-64 spp, each job on its own soup of 1k-500k triangles, c1 materials and
sky) submitted at once to the multi-job engine; every job's image must be
bit-identical to its own plain render.

Config 4: Poisson arrivals (40 jobs/s for 2 s) drawn from the config-3 size
distribution on 31 resident soups, 10% of them 1920x1080x16 spp jobs on the
config-1 scene, 32 scenes resident in all; served by the continuous-batching
loop and by the concurrent two-lane loop (small-job lane one job at a time,
and continuously batched), each image must equal the same request rendered
alone."""
import numpy as np
import pytest
import torch

from mambatracer_amd import api, serve
from mambatracer_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sky(ctx):
    dmats = ctx.to_device(W.c1_materials())
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    return dmats, env


def test_c3_all_jobs_batched_equal_plain(ctx, sky):
    print("[case] c3: 256 jobs batched vs plain", flush=True)
    dmats, env = sky
    specs, ntris = W.c3_specs()
    reqs = []
    for spec, n in zip(specs, ntris):
        sc = api.Scene(ctx, *W.soup_scene(n, spec.seed))
        reqs.append(serve.Request(0.0, sc, dmats, env, W.soup_job(spec), tag=spec.index))
    print(f"{len(reqs)} jobs, {sum(ntris)} triangles, "
          f"{sum(r.job.width * r.job.height * r.job.spp for r in reqs)} paths", flush=True)
    bat, st = serve.serve_batched(ctx, reqs, queue_capacity=1 << 23, keep_images=True)
    assert st.jobs_completed == len(reqs) and st.arena_used == 0
    got = {d.tag: d.image for d in bat}
    bad = []
    for r in reqs:
        ref, _ = api.render(ctx, r.scene, dmats, env, r.job, stats=False)
        if not torch.equal(ref, got[r.tag]):
            bad.append(r.tag)
    print(f"images differing: {bad}", flush=True)
    assert not bad


def _c4_requests(ctx, dmats, env, rate=40.0, duration=2.0):
    sizes = W.c4_soup_sizes()
    scenes = [api.Scene(ctx, *W.soup_scene(n, 5_000 + i)) for i, n in enumerate(sizes)]
    points, normals, faces, nodes, bounds, depth = W.c1_scene()
    scenes.append(api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals))
    reqs = []
    for t, spec in W.c4_arrivals(rate, duration, n_soups=len(sizes)):
        job = W.soup_job(spec) if spec.side > 0 else W.RenderJob(seed=spec.seed)
        reqs.append(serve.Request(t, scenes[spec.scene], dmats, env, job, tag=spec.index))
    return scenes, reqs


def test_c4_served_images_equal_alone(ctx, sky):
    print("[case] c4: Poisson arrivals, batched and two-lane serving vs alone", flush=True)
    dmats, env = sky
    scenes, reqs = _c4_requests(ctx, dmats, env)
    assert len(scenes) == 32
    n_c1 = sum(1 for r in reqs if r.job.width == 1920)
    print(f"{len(reqs)} requests, {n_c1} on the c1 scene", flush=True)
    assert len(reqs) >= 60 and n_c1 >= 3
    alone = {}
    for r in reqs:
        img, _ = api.render(ctx, r.scene, dmats, env, r.job, stats=False)
        alone[r.tag] = img
    bat, _ = serve.serve_batched(ctx, reqs, queue_capacity=1 << 22, keep_images=True, policy=1)
    two, _ = serve.serve_two_lane(ctx, reqs, keep_images=True)
    twob, _ = serve.serve_two_lane(ctx, reqs, keep_images=True, small_order="batched")
    for name, done in (("batched", bat), ("two-lane", two), ("two-lane batched", twob)):
        assert sorted(d.tag for d in done) == sorted(alone)
        bad = [d.tag for d in done if not torch.equal(d.image, alone[d.tag])]
        print(f"{name}: images differing {bad}", flush=True)
        assert not bad
