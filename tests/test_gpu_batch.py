"""The multi-job wavefront engine (mt_batch) against plain one-job renders.

A path runs the same per-path code under both drivers and its RNG is keyed
by (job seed, path id, bounce, purpose), so the bar is bit-exact: every
job's image from the shared queues equals its own mt_render image, whether
the job runs alone or mixed with others, whatever the queue capacity, arena
size and admission order.  (mt_render itself is held to the oracle in
test_gpu_render.py.)"""
import numpy as np
import pytest
import torch

from mambatracer_amd import _lib, api, serve
from mambatracer_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(ctx):
    """Three soups of different sizes, c1 materials plus one emitter, the c1 sky."""
    mats = W.c1_materials()
    emit = W.material_records(np.full((1, 3), 0.5), 0.04, 0.5, 0.5, emission=np.full((1, 3), 1.0),
                              emission_power=20.0)
    mats = np.concatenate([mats, emit])
    scenes = []
    for n, seed in [(1500, 41), (9000, 42), (40000, 43)]:
        points, faces, nodes, bounds, depth = W.soup_scene(n, seed)
        scenes.append(api.Scene(ctx, points, faces, nodes, bounds, depth))
    # a fourth scene whose triangles are partly emitters (mesh lights)
    points, faces, nodes, bounds, depth = W.soup_scene(5000, 44)
    faces[::50, 9] = mats.shape[0] - 1
    scenes.append(api.Scene(ctx, points, faces, nodes, bounds, depth))
    dm = ctx.to_device(mats)
    lf, nl = api.mesh_lights(ctx, scenes[3], dm)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    torch.cuda.synchronize()
    assert int(nl.item()) > 10
    return scenes, dm, env, (lf, nl)


def _jobs():
    specs = [(0, 40, 3), (1, 64, 2), (2, 33, 5), (3, 48, 2), (0, 17, 1), (2, 80, 1), (3, 24, 4)]
    out = []
    for i, (sc, side, spp) in enumerate(specs):
        j = W.soup_job(W.JobSpec(i, sc, side, spp, 9000 + i))
        if i == 6:
            j.flags = _lib.MT_RENDER_PICKED_LIGHT
        out.append((sc, j))
    return out


def _plain(ctx, world, sc, job):
    scenes, dm, env, (lf, nl) = world
    lights = (lf, nl) if sc == 3 else (None, None)
    img, _ = api.render(ctx, scenes[sc], dm, env, job, light_faces=lights[0], num_lights=lights[1],
                        stats=False)
    return img.cpu().numpy()


def _batched(ctx, world, jobs, capacity, arena, max_jobs=64, policy=_lib.MT_BATCH_OLDEST_FIRST):
    scenes, dm, env, (lf, nl) = world
    b = api.Batch(ctx, queue_capacity=capacity, arena_bytes=arena, max_jobs=max_jobs, policy=policy)
    imgs = {}
    for tag, (sc, job) in jobs:
        imgs[tag] = ctx.empty((job.width * job.height, 4), torch.float32)
        lights = (lf, nl) if sc == 3 else (None, None)
        b.submit(tag, scenes[sc], dm, env, job, imgs[tag], light_faces=lights[0], num_lights=lights[1])
    order = b.run()
    st = b.stats()
    b.close()
    torch.cuda.synchronize()
    return {t: v.cpu().numpy() for t, v in imgs.items()}, order, st


@pytest.mark.parametrize("capacity", [5000, 1 << 16])
def test_batch_matches_plain_renders_bit_exact(ctx, world, capacity):
    print(f"[case] batch mix capacity={capacity}", flush=True)
    jobs = list(enumerate(_jobs()))
    got, order, st = _batched(ctx, world, jobs, capacity, 256 << 20)
    assert sorted(order) == list(range(len(jobs)))
    assert st.jobs_completed == len(jobs) and st.active_jobs == 0 and st.pending_jobs == 0
    assert st.arena_used == 0
    rays = 0
    for tag, (sc, job) in jobs:
        ref = _plain(ctx, world, sc, job)
        np.testing.assert_array_equal(got[tag], ref, err_msg=f"job {tag}")
        assert np.isfinite(ref).all() and ref[:, :3].mean() > 0
        rays += job.width * job.height * job.spp
    assert st.extension_rays >= rays          # at least one ray per path
    if capacity == 5000:
        assert st.iterations > 10             # the queue refilled many times


def test_job_alone_equals_job_in_mix(ctx, world):
    print("[case] alone vs mix", flush=True)
    jobs = list(enumerate(_jobs()))
    mix, _, _ = _batched(ctx, world, jobs, 1 << 14, 256 << 20)
    for tag in (1, 5):
        alone, _, _ = _batched(ctx, world, [jobs[tag]], 1 << 14, 256 << 20)
        np.testing.assert_array_equal(alone[tag], mix[tag])


def test_admission_by_arena_footprint(ctx, world):
    """An arena that holds one job at a time: jobs are admitted in order as
    space frees up, complete in submission order, and images are unchanged."""
    print("[case] arena admission", flush=True)
    jobs = list(enumerate(_jobs()))[:4]
    need = [api.Batch.job_bytes(j) for _, (_, j) in jobs]
    arena = max(need)
    got, order, st = _batched(ctx, world, jobs, 1 << 14, arena)
    assert order == [0, 1, 2, 3]
    assert st.arena_peak <= arena
    for tag, (sc, job) in jobs:
        np.testing.assert_array_equal(got[tag], _plain(ctx, world, sc, job))


def test_job_slots_reused(ctx, world):
    """max_jobs = 2 slots for 7 jobs: slots and arena ranges are recycled."""
    print("[case] slot reuse", flush=True)
    jobs = list(enumerate(_jobs()))
    got, order, st = _batched(ctx, world, jobs, 1 << 15, 256 << 20, max_jobs=2)
    assert sorted(order) == list(range(len(jobs)))
    for tag in (0, 4, 6):
        sc, job = jobs[tag][1]
        np.testing.assert_array_equal(got[tag], _plain(ctx, world, sc, job))


def test_smallest_first_policy(ctx, world):
    """A short job queued behind a long one finishes first under
    MT_BATCH_SMALLEST_FIRST; images do not depend on the policy."""
    print("[case] smallest first", flush=True)
    big = W.soup_job(W.JobSpec(0, 2, 96, 4, 7001))      # 36 864 paths, 9 queue fills
    small = W.soup_job(W.JobSpec(1, 0, 16, 1, 7002))    # 256 paths
    jobs = [(0, (2, big)), (1, (0, small))]
    got, order, _ = _batched(ctx, world, jobs, 4096, 64 << 20, policy=_lib.MT_BATCH_SMALLEST_FIRST)
    assert order == [1, 0]
    np.testing.assert_array_equal(got[0], _plain(ctx, world, 2, big))
    np.testing.assert_array_equal(got[1], _plain(ctx, world, 0, small))


def test_batch_rejects_bad_jobs(ctx, world):
    scenes, dm, env, _ = world
    b = api.Batch(ctx, queue_capacity=1024, arena_bytes=1 << 20, max_jobs=4)
    img = ctx.empty((64 * 64, 4), torch.float32)
    big = W.soup_job(W.JobSpec(0, 0, 512, 4, 1))           # 16 MB of radiance > 1 MB arena
    with pytest.raises(_lib.MambartError) as e:
        b.submit(0, scenes[0], dm, env, big, ctx.empty((512 * 512, 4), torch.float32))
    assert e.value.code == -5
    bad = W.soup_job(W.JobSpec(1, 0, 64, 1, 1))
    bad.mode = _lib.MT_TRACE_REFERENCE
    with pytest.raises(_lib.MambartError):
        b.submit(1, scenes[0], dm, env, bad, img)
    bad = W.soup_job(W.JobSpec(2, 0, 64, 1, 1))
    bad.max_bounces = 64
    with pytest.raises(_lib.MambartError):
        b.submit(2, scenes[0], dm, env, bad, img)
    tags, busy = b.step()                                    # nothing queued: idle
    assert tags == [] and not busy
    with pytest.raises(_lib.MambartError):
        _lib.call("mt_batch_set_policy", b.handle, 7)
    b.close()


def test_serving_loops_agree(ctx, world):
    """The request loops (mambatracer_amd.serve): batched, first-come and
    shortest-job-first one-at-a-time serving of the same arrivals return the
    same images; shortest-job-first starts the smaller of two waiting jobs
    first."""
    from mambatracer_amd import serve
    print("[case] serving loops", flush=True)
    scenes, dm, env, (lf, nl) = world
    reqs = []
    for tag, (sc, job) in enumerate(_jobs()[:6]):
        reqs.append(serve.Request(0.0 if tag < 3 else 1.0 + 0.01 * tag, scenes[sc], dm, env, job,
                                  *((lf, nl) if sc == 3 else (None, None)), tag=tag))
    fifo, _ = serve.serve_sequential(ctx, reqs, keep_images=True)
    sjf, _ = serve.serve_sequential(ctx, reqs, keep_images=True, order="sjf")
    bat, _ = serve.serve_batched(ctx, reqs, queue_capacity=1 << 15, keep_images=True)
    ref = {d.tag: d.image.cpu().numpy() for d in fifo}
    for done in (sjf, bat):
        assert sorted(d.tag for d in done) == list(range(6))
        for d in done:
            np.testing.assert_array_equal(d.image.cpu().numpy().view(np.int32), ref[d.tag].view(np.int32))
    # the first three arrive together: shortest job first runs them smallest first
    first3 = [d.tag for d in sjf][:3]
    sizes = {r.tag: r.job.width * r.job.height * r.job.spp for r in reqs}
    assert first3 == sorted([0, 1, 2], key=lambda t: (sizes[t], t))
    with pytest.raises(ValueError):
        serve.serve_sequential(ctx, reqs[:1], order="lifo")


def test_cancel_pending_and_running_jobs(ctx, world):
    """mt_batch_cancel: a job cancelled before admission and one cancelled
    while its paths are in the queues are both reported once drained, with
    their images untouched; every other job's image stays bit-identical to
    its plain render, and the freed slot and arena serve a later job."""
    print("[case] cancel", flush=True)
    scenes, dm, env, (lf, nl) = world
    jobs = list(enumerate(_jobs()))[:5]
    b = api.Batch(ctx, queue_capacity=4096, arena_bytes=64 << 20, max_jobs=2)
    imgs = {}
    for tag, (sc, job) in jobs:
        imgs[tag] = torch.full((job.width * job.height, 4), -7.0, dtype=torch.float32, device="cuda")
        lights = (lf, nl) if sc == 3 else (None, None)
        b.submit(tag, scenes[sc], dm, env, job, imgs[tag], light_faces=lights[0], num_lights=lights[1])
    assert b.cancel(4)                     # still pending (2 slots): dropped at once
    tags, busy = b.step()                  # admits jobs 0 and 1, one iteration
    assert b.cancel(1)                     # running: drains
    assert not b.cancel(99)
    order = list(tags)
    while busy:
        t, busy = b.step()
        order.extend(t)
    # a job submitted after the cancellations reuses the freed slot / arena
    late_sc, j2 = jobs[2][1]
    late = W.soup_job(W.JobSpec(9, late_sc, j2.width, j2.spp, 9100))
    limg = ctx.empty((late.width * late.height, 4), torch.float32)
    b.submit(9, scenes[late_sc], dm, env, late, limg)
    order.extend(b.run())
    st = b.stats()
    b.close()
    torch.cuda.synchronize()
    assert sorted(order) == [0, 1, 2, 3, 4, 9]
    assert st.jobs_cancelled == 2 and st.jobs_completed == 4 and st.arena_used == 0
    for tag in (1, 4):
        assert (imgs[tag] == -7.0).all(), f"cancelled job {tag} wrote its image"
    for tag in (0, 2, 3):
        sc, job = jobs[tag][1]
        np.testing.assert_array_equal(imgs[tag].cpu().numpy(), _plain(ctx, world, sc, job), err_msg=f"job {tag}")
    np.testing.assert_array_equal(limg.cpu().numpy(), _plain(ctx, world, late_sc, late))


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("iters", [1, 3, 16])
def test_iterations_per_step(ctx, world, iters, graphs):
    """Several wavefront iterations per host round trip, replayed as a
    captured hipGraph (one per start parity of the queue ping-pong: an odd
    count alternates them) or launched kernel by kernel: same images."""
    print(f"[case] {iters} iterations per step, graphs={graphs}", flush=True)
    scenes, dm, env, (lf, nl) = world
    jobs = list(enumerate(_jobs()))
    b = api.Batch(ctx, queue_capacity=1 << 14, arena_bytes=256 << 20, max_jobs=64)
    b.set_graphs(graphs)
    b.set_iterations(iters)
    imgs = {}
    for tag, (sc, job) in jobs:
        imgs[tag] = ctx.empty((job.width * job.height, 4), torch.float32)
        lights = (lf, nl) if sc == 3 else (None, None)
        b.submit(tag, scenes[sc], dm, env, job, imgs[tag], light_faces=lights[0], num_lights=lights[1])
    order = b.run()
    b.close()
    assert sorted(order) == list(range(len(jobs)))
    for tag, (sc, job) in jobs:
        np.testing.assert_array_equal(imgs[tag].cpu().numpy(), _plain(ctx, world, sc, job), err_msg=f"job {tag}")


def test_graph_recaptured_when_iterations_change(ctx, world):
    """Changing the iterations per step between steps, with jobs in flight,
    records new graphs; every image still equals its plain render."""
    print("[case] iterations change mid-run (graphs)", flush=True)
    scenes, dm, env, (lf, nl) = world
    jobs = list(enumerate(_jobs()))
    b = api.Batch(ctx, queue_capacity=1 << 13, arena_bytes=256 << 20, max_jobs=64)
    imgs = {}
    for tag, (sc, job) in jobs:
        imgs[tag] = ctx.empty((job.width * job.height, 4), torch.float32)
        lights = (lf, nl) if sc == 3 else (None, None)
        b.submit(tag, scenes[sc], dm, env, job, imgs[tag], light_faces=lights[0], num_lights=lights[1])
    done = []
    for iters in (2, 5, 1, 4):
        b.set_iterations(iters)
        for _ in range(3):
            tags, busy = b.step()
            done += tags
    done += b.run()
    b.close()
    assert sorted(done) == list(range(len(jobs)))
    for tag, (sc, job) in jobs:
        np.testing.assert_array_equal(imgs[tag].cpu().numpy(), _plain(ctx, world, sc, job), err_msg=f"job {tag}")


@pytest.mark.parametrize("mode", [_lib.MT_TRACE_PACKED_STACK, _lib.MT_TRACE_PACKED_POOL, _lib.MT_BATCH_TRACE_SPLIT])
def test_pool_traversal_in_the_engine(ctx, world, mode):
    """The shared queue traced one ray per lane, with wave pools (each lane
    loading its slot's scene), or split (pools for the continuing paths, one
    ray per lane for the new camera rays; the default) gives the images of
    plain renders."""
    print(f"[case] engine with traversal mode {mode}", flush=True)
    scenes, dm, env, (lf, nl) = world
    jobs = list(enumerate(_jobs()))
    b = api.Batch(ctx, queue_capacity=1 << 14, arena_bytes=256 << 20, max_jobs=64)
    b.set_trace_mode(mode)
    b.set_iterations(3)
    imgs = {}
    for tag, (sc, job) in jobs:
        imgs[tag] = ctx.empty((job.width * job.height, 4), torch.float32)
        lights = (lf, nl) if sc == 3 else (None, None)
        b.submit(tag, scenes[sc], dm, env, job, imgs[tag], light_faces=lights[0], num_lights=lights[1])
    order = b.run()
    b.close()
    assert sorted(order) == list(range(len(jobs)))
    for tag, (sc, job) in jobs:
        np.testing.assert_array_equal(imgs[tag].cpu().numpy(), _plain(ctx, world, sc, job), err_msg=f"job {tag}")


def test_engine_autotuner(ctx, world):
    """serve.autotune_batch tries each setting dimension in turn on a sample
    and returns one of the tried settings; serving with it reproduces the
    plain images."""
    print("[case] engine autotuner", flush=True)
    scenes, dm, env, (lf, nl) = world
    jobs = [(t, (sc, job)) for t, (sc, job) in enumerate(_jobs()) if sc != 3]
    reqs = [serve.Request(0.0, scenes[sc], dm, env, job, tag=t) for t, (sc, job) in jobs]
    best, trials = serve.autotune_batch(ctx, reqs, arena_bytes=256 << 20, max_jobs=64,
                                        queue_capacities=(1 << 12, 1 << 14), iterations=(1, 4))
    assert set(best) == {"queue_capacity", "iterations", "trace_mode", "graphs"}
    assert any(t == best for t, _ in trials) and all(r > 0 for _, r in trials)
    assert len(trials) >= 5
    b = serve.make_batch(ctx, best, arena_bytes=256 << 20, max_jobs=64)
    done, _ = serve.serve_batched(ctx, reqs, keep_images=True, batch=b)
    b.close()
    for d in done:
        sc, job = dict(jobs)[d.tag]
        np.testing.assert_array_equal(d.image.cpu().numpy(), _plain(ctx, world, sc, job), err_msg=f"job {d.tag}")
