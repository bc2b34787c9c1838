"""Render-request loop over the multi-job wavefront engine (continuous
batching), and the one-job-at-a-time loop it is measured against.

The reference has one render job per device at a time (CDeviceGI.start ->
renderThread -> execute -> loop, src/cl/device/CDeviceGI.java:190-204,
424-431).  `serve_batched` accepts jobs as they arrive, submits them to an
`api.Batch` (admission by arena footprint happens in the engine, in arrival
order), keeps stepping the shared queues and timestamps each job when its
image is complete.  `serve_sequential` runs the same arrivals through plain
`api.render` calls, first come first served or shortest job first.
`serve_two_lane` runs two lanes on two streams of the device, one host
thread each: jobs that fill the GPU by themselves go to one lane (plain
renders, first come first served), the rest to the other, so a small job
never waits behind a large one's whole render; the small-job lane renders
one job at a time shortest first, or runs its jobs through a multi-job
engine of its own (continuous batching: every waiting small job is admitted
at once, whatever the large lane is doing).

A request is a `Request`: arrival time (seconds from the start of the run),
the resident scene, materials, environment and lights it renders with, and
the RenderJob.  Both loops return one `Done` record per request.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

from . import api


@dataclass
class Request:
    arrival: float
    scene: Any
    materials: Any
    env: Any
    job: Any
    light_faces: Any = None
    num_lights: Any = None
    tag: int = 0


@dataclass
class Done:
    tag: int
    arrival: float
    complete: float
    image: Any = field(repr=False, default=None)

    @property
    def latency(self) -> float:
        return self.complete - self.arrival


def _wait_until(t0: float, t: float) -> None:
    while True:
        dt = t - (time.perf_counter() - t0)
        if dt <= 0:
            return
        time.sleep(min(dt, 0.0005))


def serve_batched(ctx: api.Context, requests: List[Request], queue_capacity: int = 1 << 23,
                  arena_bytes: int = 32 << 30, max_jobs: int = 1024, keep_images: bool = False,
                  batch: Optional[api.Batch] = None, policy: int = 0):
    """Continuous batching: returns (Done records in completion order, engine stats).
    policy: MT_BATCH_OLDEST_FIRST (0) or MT_BATCH_SMALLEST_FIRST (1)."""
    torch = api._torch()
    reqs = sorted(requests, key=lambda r: r.arrival)
    own = batch is None
    b = api.Batch(ctx, queue_capacity, arena_bytes, max_jobs, policy=policy) if own else batch
    images, arrivals = {}, {}
    out: List[Done] = []
    t0 = time.perf_counter()
    i, busy = 0, False
    while i < len(reqs) or busy:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i].arrival <= now:
            r = reqs[i]
            img = ctx.empty((r.job.width * r.job.height, 4), torch.float32)
            b.submit(r.tag, r.scene, r.materials, r.env, r.job, img, r.light_faces, r.num_lights)
            images[r.tag], arrivals[r.tag] = img, r.arrival
            i += 1
            busy = True
        if not busy:
            _wait_until(t0, reqs[i].arrival)
            continue
        tags, busy = b.step()
        t = time.perf_counter() - t0
        for tag in tags:
            img = images.pop(tag)
            out.append(Done(tag, arrivals.pop(tag), t, img if keep_images else None))
    stats = b.stats()
    if own:
        b.close()
    return out, stats


def _paths(r: Request) -> int:
    return r.job.width * r.job.height * r.job.spp


def serve_sequential(ctx: api.Context, requests: List[Request], keep_images: bool = False, order: str = "fifo"):
    """One job at a time with plain mt_render launches.  order "fifo": first
    come first served; "sjf": whenever the GPU frees up, the waiting job with
    the fewest paths runs next (ties by arrival)."""
    if order not in ("fifo", "sjf"):
        raise ValueError(f"unknown order {order!r}")
    reqs = sorted(requests, key=lambda r: r.arrival)
    out: List[Done] = []
    rays = 0
    waiting: List[Request] = []
    i = 0
    t0 = time.perf_counter()
    while i < len(reqs) or waiting:
        if not waiting:
            _wait_until(t0, reqs[i].arrival)
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i].arrival <= now:
            waiting.append(reqs[i])
            i += 1
        if order == "sjf":
            k = min(range(len(waiting)), key=lambda j: (_paths(waiting[j]), waiting[j].arrival))
        else:
            k = 0
        r = waiting.pop(k)
        img, st = api.render(ctx, r.scene, r.materials, r.env, r.job, light_faces=r.light_faces,
                             num_lights=r.num_lights, stats=True)   # stats read-back synchronises
        rays += int(st.extension_rays) + int(st.shadow_rays)
        out.append(Done(r.tag, r.arrival, time.perf_counter() - t0, img if keep_images else None))
    return out, rays


def serve_two_lane(ctx: api.Context, requests: List[Request], keep_images: bool = False,
                   big_paths: int = 1 << 22, small_order: str = "sjf", small_queue: int = 1 << 21,
                   small_arena: int = 16 << 30, small_iterations: int = 1):
    """Two render lanes, each a context on its own stream of ctx's device
    driven by its own host thread: requests of at least `big_paths` paths
    (a 1080p x 16 spp job, a 1024^2 x 64 spp soup) go to the large lane,
    first come first served; the rest to the small lane: one plain render
    at a time, shortest job first ("sjf") or in arrival order ("fifo"), or
    ("batched") a multi-job engine of the lane's own (queue capacity
    `small_queue` paths, arena `small_arena` bytes, smallest job first,
    `small_iterations` bounce iterations per step) that admits every waiting
    small job at once.  The lanes' kernels share the GPU, so small jobs are
    not queued behind a large job's whole render.  Plain renders and the
    engine give each job the image it has alone.  Returns (Done records in
    completion order, rays traced)."""
    if small_order not in ("fifo", "sjf", "batched"):
        raise ValueError(f"unknown order {small_order!r}")
    torch = api._torch()
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=ctx.torch_device), torch.cuda.Stream(device=ctx.torch_device)]
    lanes_ctx = [api.Context(ctx.device, stream=s) for s in streams]
    reqs = sorted(requests, key=lambda r: r.arrival)
    waiting: List[List[Request]] = [[], []]
    cond = threading.Condition()
    closed = [False]
    out: List[Done] = []
    rays = [0]
    errors: List[BaseException] = []
    t0 = time.perf_counter()

    def batched_lane(k: int) -> None:
        torch_ = api._torch()
        lctx = lanes_ctx[k]
        b = None
        try:
            b = api.Batch(lctx, small_queue, small_arena, 1024, policy=1)
            b.set_iterations(small_iterations)
            images, arrivals = {}, {}
            busy = False
            while True:
                with cond:
                    while not waiting[k] and not closed[0] and not busy:
                        cond.wait()
                    if not waiting[k] and closed[0] and not busy:
                        break
                    new, waiting[k][:] = waiting[k][:], []
                for r in new:
                    img = lctx.empty((r.job.width * r.job.height, 4), torch_.float32)
                    b.submit(r.tag, r.scene, r.materials, r.env, r.job, img, r.light_faces, r.num_lights)
                    images[r.tag], arrivals[r.tag] = img, r.arrival
                    busy = True
                if busy:
                    tags, busy = b.step()
                    t = time.perf_counter() - t0
                    with cond:
                        for tag in tags:
                            img = images.pop(tag)
                            out.append(Done(tag, arrivals.pop(tag), t, img if keep_images else None))
            st = b.stats()
            with cond:
                rays[0] += int(st.extension_rays) + int(st.shadow_rays)
        except BaseException as e:  # reported by the caller's thread
            errors.append(e)
            with cond:  # nothing more is accepted here: let the feeder finish
                closed[0] = True
        finally:
            if b is not None:
                b.close()

    def lane(k: int) -> None:
        if k == 0 and small_order == "batched":
            batched_lane(k)
            return
        try:
            while True:
                with cond:
                    while not waiting[k] and not closed[0]:
                        cond.wait()
                    if not waiting[k]:
                        return
                    q = waiting[k]
                    i = 0
                    if k == 0 and small_order == "sjf":
                        i = min(range(len(q)), key=lambda j: (_paths(q[j]), q[j].arrival))
                    r = q.pop(i)
                img, st = api.render(lanes_ctx[k], r.scene, r.materials, r.env, r.job, light_faces=r.light_faces,
                                     num_lights=r.num_lights, stats=True)   # synchronises this lane
                t = time.perf_counter() - t0
                with cond:
                    rays[0] += int(st.extension_rays) + int(st.shadow_rays)
                    out.append(Done(r.tag, r.arrival, t, img if keep_images else None))
        except BaseException as e:  # reported by the caller's thread
            errors.append(e)

    threads = [threading.Thread(target=lane, args=(k,), daemon=True) for k in (0, 1)]
    for th in threads:
        th.start()
    try:
        for r in reqs:
            _wait_until(t0, r.arrival)
            with cond:
                waiting[1 if _paths(r) >= big_paths else 0].append(r)
                cond.notify_all()
    finally:
        with cond:
            closed[0] = True
            cond.notify_all()
        for th in threads:
            th.join()
        for c in lanes_ctx:
            c.close()
    if errors:
        raise errors[0]
    out.sort(key=lambda d: d.complete)
    return out, rays[0]


def make_batch(ctx: api.Context, setting: dict, arena_bytes: int = 32 << 30, max_jobs: int = 1024,
               policy: int = 1) -> api.Batch:
    """A multi-job engine with an autotuner setting: queue_capacity,
    iterations (per step), trace_mode (_lib.MT_TRACE_PACKED_STACK / _POOL /
    _lib.MT_BATCH_TRACE_SPLIT),
    graphs (replay captured iterations)."""
    b = api.Batch(ctx, int(setting["queue_capacity"]), arena_bytes, max_jobs, policy=policy)
    b.set_iterations(int(setting["iterations"]))
    b.set_trace_mode(int(setting["trace_mode"]))
    b.set_graphs(bool(setting["graphs"]))
    return b


def autotune_batch(ctx: api.Context, requests: List[Request], arena_bytes: int = 32 << 30,
                   max_jobs: int = 1024, policy: int = 1,
                   queue_capacities=(1 << 21, 1 << 22, 1 << 23), iterations=(1, 4, 8, 16),
                   trace_modes=None, graphs=(True, False), objective: str = "jobs_per_s"):
    """Chooses the multi-job engine's settings for a workload: the queue
    capacity (paths in flight per iteration), the bounce iterations per host
    round trip, the closest-hit traversal of the shared queue (one ray per
    lane, wave pools, or pools for the continuing paths and one ray per lane
    for the new ones) and graph replay.  `requests` is a sample of the
    workload (its scenes, sizes and sample counts), served through
    serve_batched per trial: objective "jobs_per_s" submits them at once and
    the setting with the most jobs/s wins; "p99" keeps their arrival times
    and the lowest p99 latency wins.  Coordinate search from (largest capacity, 8 iterations, the
    split traversal, graphs): each dimension in turn takes its best value with
    the others held.  Every setting renders the same images.  Returns (best
    setting, [(setting, jobs/s or p99 ms), ...])."""
    from . import _lib
    if objective not in ("jobs_per_s", "p99"):
        raise ValueError(f"unknown objective {objective!r}")
    if trace_modes is None:
        trace_modes = (_lib.MT_BATCH_TRACE_SPLIT, _lib.MT_TRACE_PACKED_STACK, _lib.MT_TRACE_PACKED_POOL)
    sample = [Request(0.0 if objective == "jobs_per_s" else r.arrival, r.scene, r.materials, r.env, r.job,
                      r.light_faces, r.num_lights, tag=k) for k, r in enumerate(requests)]
    best = {"queue_capacity": max(queue_capacities), "iterations": 8 if 8 in iterations else iterations[0],
            "trace_mode": trace_modes[0], "graphs": graphs[0]}
    trials = []
    scores = {}

    def score(setting):
        key = tuple(sorted(setting.items()))
        if key in scores:
            return scores[key]
        b = make_batch(ctx, setting, arena_bytes, max_jobs, policy)
        try:
            done, _ = serve_batched(ctx, sample, batch=b)
        finally:
            b.close()
        sm = summary(done, 0)
        rate = sm["jobs_per_s"] if objective == "jobs_per_s" else -sm["latency_ms"]["p99"]
        scores[key] = rate
        trials.append((dict(setting), sm["jobs_per_s"] if objective == "jobs_per_s" else sm["latency_ms"]["p99"]))
        return rate

    score(best)  # warm-up trial (code objects, allocator), not compared
    scores.clear()
    trials.clear()
    for name, values in (("queue_capacity", queue_capacities), ("iterations", iterations),
                         ("trace_mode", trace_modes), ("graphs", graphs)):
        rates = {v: score({**best, name: v}) for v in values}
        best[name] = max(rates, key=lambda v: rates[v])
    return best, trials


def summary(done: List[Done], rays: int) -> dict:
    """Latency percentiles (ms), jobs/s and Mrays/s over the run's span."""
    lat = np.array([d.latency for d in done]) * 1e3
    span = max(d.complete for d in done) - min(d.arrival for d in done)
    return {
        "jobs": len(done),
        "span_s": round(span, 4),
        "jobs_per_s": round(len(done) / span, 3),
        "mrays_per_s": round(rays / span / 1e6, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
    }
