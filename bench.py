#!/usr/bin/env python3
"""Throughput of the wavefront path-tracing core on one MI355X.

Workload (one "step" = one render job, inputs resident in HBM): 256 UV
spheres x 4096 triangles + ground quad (1,048,578 triangles), GGX materials,
1024x512 HDR environment sampled through its summed-area table, 1920x1080 at
16 spp, up to 8 bounces with Russian roulette after bounce 3; after every bounce the
surviving paths form the next bounce's queue.  value = (closest-hit + shadow rays) / s over
the timed steps, each one replay of the job's captured render
(mt_render_graph, as a render thread re-runs a frame) plus its ray-count
readback.  The same job through uncaptured mt_render is timed after it and
reported beside it; both produce the same image bit for bit (checked).

roofline: the dominant kernel is the closest-hit traversal.  Its algorithmic
bytes are the reference layout's per-ray figure 40 + 148*I + 116*L + 28*P
(I interior steps, L leaf visits, P backtrack pops, counted by the device's
counting traversal in an untimed render of the same job) divided by the
traversal launches' summed duration (HIP events on the launch stream around
every closest-hit launch of the uncaptured renders).

cpu_baseline: the oracle (plain-C restatement, test infrastructure) traces a
bounded random subset of the same job's paths on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s path-traced (1M tris, 1080p) + p50/p99 render-job latency at fixed QPS"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(workload: str = "c1", mode=None):
    """HBM bytes per closest-hit launch measured by rocprofv3 --pmc (profiles/), for the
    traversal mode `mode` where the profile holds one figure per mode; or None."""
    p = os.path.join(ROOT, "profiles", f"traffic_trace_closest_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        if mode is not None and "modes" in d:
            m = d["modes"].get(str(int(mode)))
            return float(m["bytes"]) if m else None
        return float(d["bytes_per_launch"])
    except Exception:
        return None


def valu_from_profiles(workload: str = "c1", mode=None):
    """VALU instructions per closest-hit launch (rocprofv3 SQ_INSTS_VALU, profiles/) and the
    issue peak they are compared with, or None (also when the profile was taken for
    another traversal mode than `mode`)."""
    p = os.path.join(ROOT, "profiles", f"counters_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        if mode is not None and "mode" in d and int(d["mode"]) != int(mode):
            return None
        return (float(d["valu_inst_per_launch"]), float(d["valu_peak_inst_per_s"]), d["valu_peak_basis"],
                d.get("cycles_per_inst"))
    except Exception:
        return None


# wave64 VALU instructions per lane and loop iteration of each calibration mix
# (mt_calib.hip; counts read from the kernels' gfx950 assembly)
VALU_CALIB_MIX = {0: ("v_mul_f32 + v_add_f32", 16), 1: ("v_mul_f32/v_add_f32 + 25% v_rcp_f32", 16),
                  2: ("v_min_f32/v_max_f32 (+ NaN-quieting v_max)", 24), 3: ("v_add_u32 + v_xor_b32", 16),
                  4: ("v_cmp_f32 + v_cndmask_b32 + v_add_f32", 24), 5: ("v_fma_f32", 16),
                  6: ("8 correctly rounded f32 divisions (11 instructions each)", 88),
                  7: ("8 correctly rounded sqrtf (16 instructions each) + v_add_f32", 136)}


def valu_calibration(blocks: int = 8192, iters: int = 4096, reps: int = 5):
    """Issue rate of wave64 VALU instructions on this GPU, per mix: blocks x
    256 lanes run `iters` iterations of independent chains (libmt_calib.so);
    the best of `reps` launches timed with HIP events on the launch stream."""
    import ctypes

    import torch
    lib = ctypes.CDLL(os.path.join(ROOT, "mambatracer_amd", "libmt_calib.so"))
    lib.mt_calib_valu.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.mt_calib_valu.restype = ctypes.c_int
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    res = {"blocks": blocks, "lanes_per_block": 256, "iters": iters, "mixes": {}}
    waves = blocks * 256 // 64
    for kind, (what, per_iter) in VALU_CALIB_MIX.items():
        best = None
        for _ in range(reps + 1):  # first launch untimed
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            rc = lib.mt_calib_valu(kind, blocks, iters, ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(stream.cuda_stream))
            e1.record(stream)
            if rc != 0:
                raise RuntimeError(f"mt_calib_valu({kind}) returned {rc}")
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None or ms < best else best
        inst = per_iter * iters * waves
        res["mixes"][str(kind)] = {"mix": what, "loop_inst": inst, "ms": round(best, 4),
                                   "G_inst_per_s": round(inst / (best * 1e-3) / 1e9, 1)}
    res["checksum"] = float(out.double().sum())
    return res


def cpu_baseline(W, scene_arrays, mats, rgba, job, npaths, seconds, threads, grid=None, grid_mode=0):
    import numpy as np
    from oracle import orc_py as orc
    points, normals, faces, nodes, bounds, _ = scene_arrays
    w, h = rgba.shape[1], rgba.shape[0]
    lum, sat = orc.env_lum_sat(rgba.reshape(-1, 4), w, h)
    env = (rgba.reshape(-1, 4), lum, sat, w, h)
    rng = np.random.default_rng(12345)
    # untimed warm-up (threads, page faults), then a timed probe sizes the sample
    warm = np.sort(rng.choice(npaths, 20000, replace=False)).astype(np.int32)
    kw = dict(nthreads=threads, grid=grid, grid_mode=grid_mode)
    orc.render_paths(points, normals, faces, nodes, bounds, mats, env, job, warm, **kw)
    probe = np.sort(rng.choice(npaths, 400000, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    orc.render_paths(points, normals, faces, nodes, bounds, mats, env, job, probe, **kw)
    dt = time.perf_counter() - t0
    n = int(min(npaths, max(400000, 400000 * seconds / max(dt, 1e-3))))
    sample = np.sort(rng.choice(npaths, n, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    _, rays = orc.render_paths(points, normals, faces, nodes, bounds, mats, env, job, sample, **kw)
    dt = time.perf_counter() - t0
    return {
        "value": round(float(rays.sum()) / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} of {npaths} path ids of the same job drawn uniformly (seed 12345), "
                  f"{int(rays[0])} closest-hit + {int(rays[1])} shadow rays in {dt:.1f} s",
    }


def cpu_baseline_c2(points, faces, nodes, bounds, o, d, light, seconds, threads):
    """The oracle's closest hit, shadow-ray setup and any hit (the same stages
    as the GPU step) on a bounded sample of c2's rays: rays/s over the two
    traversals, scaled to Grays/s."""
    import numpy as np
    from oracle import orc_py as orc
    rng = np.random.default_rng(12345)

    def run(idx):
        oo, dd = o[idx], d[idx]
        prim, t, _, _ = orc.trace_closest(points, faces, nodes, bounds, oo, dd, nthreads=threads)
        so, sd, _ = orc.shadow_rays(oo, dd, prim, t, light)
        orc.trace_any(points, faces, nodes, bounds, so, sd, nthreads=threads)

    n_all = len(o)
    run(np.sort(rng.choice(n_all, 2000, replace=False)))          # untimed warm-up
    probe = np.sort(rng.choice(n_all, 50000, replace=False))
    t0 = time.perf_counter()
    run(probe)
    dt = time.perf_counter() - t0
    n = int(min(n_all, max(50000, 50000 * seconds / max(dt, 1e-3))))
    sample = np.sort(rng.choice(n_all, n, replace=False))
    t0 = time.perf_counter()
    run(sample)
    dt = time.perf_counter() - t0
    return {
        "value": round(2 * n / dt / 1e9, 5),
        "unit": "Grays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} of {n_all} rays drawn uniformly (seed 12345): closest hit, shadow rays to the light, "
                  f"any hit, {2 * n} rays in {dt:.1f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--mode", type=int, default=0, help="0 packed traversal layout, 1 reference layout")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--lighting", choices=["guided", "guided-tile-pdf", "full-map"], default="guided",
                    help="c1 environment lighting: the adaptive light grid with the reference's pdf (default), "
                         "the same with the tile pick's pdf applied, or the full-map SAT")
    ap.add_argument("--qps-seconds", type=float, default=3.0,
                    help="c1: seconds of Poisson arrivals at 70%% of the measured capacity for p50/p99 latency")
    ap.add_argument("--other-steps", type=int, default=5,
                    help="c1: timed replays of each other lighting path, reported beside the headline")
    ap.add_argument("--tail", type=int, default=3,
                    help="c1: bounces from this one on run in one launch, each lane carrying a path to its end "
                         "(MT_RENDER_TAIL_FROM); 0: three launches per bounce")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "compact", "bvh", "stages"], default="c1",
                    help="c1: one 1080p render per step (default line); c2: 64M random rays through 4M "
                         "triangles (closest hit, shadow rays, any hit, compaction); c3: 256 uneven jobs "
                         "batched vs one at a time; c4: Poisson-arrival serving loop; compact: stable "
                         "compaction alone on the shapes the paths use; bvh: the device BVH builder on the "
                         "c1 and c2 scenes; stages: BSDF, light-grid and environment kernels alone")
    ap.add_argument("--rays", type=int, default=64 << 20, help="c2: rays per step")
    ap.add_argument("--jobs", type=int, default=256, help="c3: number of jobs")
    ap.add_argument("--rate", type=float, default=40.0, help="c4: arrivals per second")
    ap.add_argument("--duration", type=float, default=5.0, help="c4: seconds of arrivals")
    ap.add_argument("--queue", type=int, default=1 << 23, help="c3/c4: wavefront queue capacity")
    ap.add_argument("--policy", type=int, default=0, help="c3/c4: 0 oldest job first, 1 smallest job first")
    ap.add_argument("--iters", type=int, default=1, help="c3/c4: wavefront iterations per batch step")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy 1 GiB device to device and exit (PMC byte-counter calibration)")
    ap.add_argument("--gather-calibration", action="store_true",
                    help="gather 1 GiB of 64-byte, 64-byte-aligned records in a random order and exit (PMC "
                         "byte-counter calibration for the traversal's dependent 64-byte node reads)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="time the vector-ALU issue-rate kernels (libmt_calib.so) and print their rates "
                         "(the measured peak of the closest-hit kernels' instruction-issue roofline)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by pixels or jobs "
              "with no exchange, see DESIGN.md.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from mambatracer_amd import _lib, api
    from mambatracer_amd import workloads as W

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied {a.numel() * 4} bytes", flush=True)
        return 0
    if args.gather_calibration:
        # 16M records of 64 B (a float32 [n, 16] tensor, rows 64-B aligned),
        # each read exactly once in a random order (index_select's gather
        # kernel), written out in order
        n = 1 << 24
        src = torch.rand((n, 16), dtype=torch.float32, device="cuda")
        perm = torch.randperm(n, device="cuda")
        torch.cuda.synchronize()
        out = torch.index_select(src, 0, perm)
        torch.cuda.synchronize()
        print(f"gathered {n} records of 64 bytes ({n * 64} bytes) in a random order; "
              f"checksum {float(out[:, 0].sum()):.3f}", flush=True)
        return 0
    if args.valu_calibration:
        print(json.dumps(valu_calibration()), flush=True)
        return 0
    if args.workload == "c2":
        return bench_c2(args)
    if args.workload == "compact":
        return bench_compact(args)
    if args.workload == "bvh":
        return bench_bvh(args)
    if args.workload == "stages":
        return bench_stages(args)
    if args.workload != "c1":
        return bench_jobs(args)
    return bench_c1(args)


def cpu_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on
    (its affinity mask), capped by OMP_NUM_THREADS when that is set (the
    share of a machine's cores a GPU job is given)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


LIGHTING = {
    "guided": ("env 1024x512 sampled via the adaptive light grid (illuminateEnvironmentLight1 / "
               "getRadianceEnvironmentLight1 + accumLightGrid, the reference's in-tile pdf) with its tile SATs"),
    "guided-tile-pdf": ("env 1024x512 sampled via the adaptive light grid with its tile SATs, tile-pick pdf "
                        "applied"),
    "full-map": "env 1024x512 sampled via its full-map SAT",
}


def lighting_flags(lighting: str) -> int:
    from mambatracer_amd import _lib
    if lighting == "full-map":
        return 0
    f = _lib.MT_RENDER_LIGHTGRID_ENV
    return f | (_lib.MT_RENDER_LIGHTGRID_TILE_PDF if lighting == "guided-tile-pdf" else 0)


def prepare_grid(api, ctx, env, rgba, job, render_frame, frames: int = 5):
    """The light grid as the reference's render thread leaves it after its
    adaptive phase: seeded from the environment (setHDRLuminance), camera
    position set (CDeviceGI.java:296), then frames 0..frames-1 each followed
    by mt_lightgrid_update (CDeviceGI.java:263-264).  Returns the grid."""
    import numpy as np
    w, h = rgba.shape[1], rgba.shape[0]
    lum = np.asarray(env.lum.cpu().numpy(), np.float32)
    grid = api.CAdaptiveEnvironment(ctx)
    grid.setHDRLuminance(api.lightgrid_seed_tile(lum, w, h))
    grid.setCameraPosition(job.pos)
    env.attachLightGrid(grid)
    for f in range(frames):
        render_frame(1000 + f)
        grid.update()
    return grid


def latency_at_qps(graph, seed, rate, duration, rng_seed=7):
    """p50/p99 of render-job latency (arrival -> image complete) for jobs of
    this view arriving as a Poisson process of `rate` per second for
    `duration` seconds, served in arrival order on the render stream: each
    job is launched at its arrival time (queued behind earlier jobs when the
    GPU is busy) and completes at its end event."""
    import numpy as np
    import torch
    rng = np.random.default_rng(rng_seed)
    t, arrivals = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        arrivals.append(t)
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ends = [torch.cuda.Event(enable_timing=True) for _ in arrivals]
    torch.cuda.synchronize()
    ev0.record(stream)
    h0 = time.perf_counter()
    for i, a in enumerate(arrivals):
        wait = a - (time.perf_counter() - h0)
        if wait > 0:
            time.sleep(wait)
        graph.launch(seed + i)
        ends[i].record(stream)
    torch.cuda.synchronize()
    lat = np.array([ev0.elapsed_time(e) for e in ends]) - np.array(arrivals) * 1e3
    lat = np.maximum(lat, 0.0)
    return {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
            "qps": round(rate, 2), "jobs": len(arrivals), "duration_s": duration,
            "mean": round(float(lat.mean()), 3)}


def bench_c1(args) -> int:
    import numpy as np
    import torch

    from mambatracer_amd import _lib, api
    from mambatracer_amd import workloads as W

    t0 = time.perf_counter()
    scene_arrays = W.c1_scene()
    points, normals, faces, nodes, bounds, depth = scene_arrays
    mats = W.c1_materials()
    rgba = W.env_map()
    log(f"scene: {faces.shape[0]} triangles, BVH depth {depth}, host prep {time.perf_counter() - t0:.1f} s")

    ctx = api.Context(0)
    scene = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    dmats = ctx.to_device(mats)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(rgba)
    job = W.RenderJob(width=args.width, height=args.height, spp=args.spp, mode=args.mode)
    npaths = job.width * job.height * job.spp
    lflags = lighting_flags(args.lighting)
    # bounces >= args.tail in one launch (k_tail): this open scene leaves few
    # paths after the first bounces
    tflags = _lib.MT_RENDER_TAIL_FROM(args.tail) if 0 < args.tail < job.max_bounces else 0

    grid = None
    if lflags:
        def frame(seed):
            job.seed, job.flags = seed, lflags | tflags
            api.render(ctx, scene, dmats, env, job, stats=False)
        grid = prepare_grid(api, ctx, env, rgba, job, frame)
        job.seed = W.RenderJob().seed
        log("light grid prepared: 5 frames, each followed by an update")

    # untimed: counting traversal for the algorithmic bytes of the
    # closest-hit launches the roofline times (bounces before the tail; the
    # counting loop runs per-bounce launches and the same paths)
    job.flags = _lib.MT_RENDER_TRAVERSAL_STATS | lflags
    nb_all = job.max_bounces
    if tflags:
        job.max_bounces = args.tail
    _, st = api.render(ctx, scene, dmats, env, job)
    job.max_bounces = nb_all
    ilp = [int(x) for x in st.closest_ilp]
    ext_rays = int(st.extension_rays)
    alg_bytes = 40 * ext_rays + 148 * ilp[0] + 116 * ilp[1] + 28 * ilp[2]
    log(f"traversal counters: rays {ext_rays}, I {ilp[0]}, L {ilp[1]}, P {ilp[2]} "
        f"({alg_bytes / max(ext_rays, 1):.0f} B/ray)")

    # The job's launch sequence captured once (mt_render_graph: state laid
    # out once, one graph launch per frame); each step replays it with the
    # job's seed and reads its ray counts.
    job.flags = lflags | tflags
    graph = api.RenderGraph(ctx, scene, dmats, env, job)
    for i in range(args.warmup):
        graph.launch(job.seed)
        graph.stats()
        log(f"warmup {i + 1}/{args.warmup} done")

    lat, rays = [], 0
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        graph.launch(job.seed)
        st = graph.stats()   # synchronises
        lat.append((time.perf_counter() - s0) * 1e3)
        rays += int(st.extension_rays) + int(st.shadow_rays)
        log(f"step {i + 1}/{args.steps}: {lat[-1]:.1f} ms, {int(st.extension_rays)} + {int(st.shadow_rays)} rays")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    img = graph.image.clone()
    bounce_paths = [int(st.bounce_counts[b]) for b in range(64) if st.bounce_counts[b] > 0]
    log(f"live paths entering each bounce: {bounce_paths}")

    # The same job through mt_render (uncaptured), with every traversal
    # launch bracketed by HIP events on the context stream: the closest-hit
    # kernel's average launch duration for the roofline, and the uncaptured
    # throughput for comparison.
    job.flags = _lib.MT_RENDER_PROFILE | lflags | tflags
    api.render(ctx, scene, dmats, env, job)
    e_rays, closest_ms, launches = 0, 0.0, 0
    torch.cuda.synchronize()
    e_start = time.perf_counter()
    for i in range(args.steps):
        eimg, est = api.render(ctx, scene, dmats, env, job)
        e_rays += int(est.extension_rays) + int(est.shadow_rays)
        closest_ms += float(est.closest_ms)
        launches += int(est.closest_launches)
    torch.cuda.synchronize()
    e_elapsed = time.perf_counter() - e_start
    if not lflags and not torch.equal(eimg, img):
        # (with the light grid, each frame's sums change the grid the next
        # frame samples, so consecutive frames differ by design)
        log("captured and uncaptured renders differ")
        return 1
    log(f"uncaptured: {e_elapsed * 1e3 / args.steps:.2f} ms per job")

    img_np = img.cpu().numpy()
    if not np.isfinite(img_np).all():
        log("non-finite pixels in the image")
        return 1
    value = rays / elapsed / 1e6
    per_launch_bytes = alg_bytes / max(est.closest_launches, 1)
    avg_launch_s = closest_ms / max(launches, 1) / 1e3
    alg_achieved = per_launch_bytes / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    traffic = traffic_from_profiles("c1")
    achieved = traffic / avg_launch_s / 1e9 if (traffic and avg_launch_s > 0) else None

    # render-job latency at a fixed arrival rate: 70% of the measured capacity
    qps = None
    if args.qps_seconds > 0:
        rate = 0.7 * args.steps / elapsed
        job.flags = lflags | tflags
        qps = latency_at_qps(graph, job.seed, rate, args.qps_seconds)
        log(f"latency at {rate:.1f} jobs/s: {qps}")

    # the other lighting paths, beside the headline (captured, same view)
    others = {}
    for other in LIGHTING:
        if other == args.lighting or args.other_steps <= 0:
            continue
        of = lighting_flags(other)
        if of and grid is None:
            def frame(seed, of=of):
                job.seed, job.flags = seed, of | tflags
                api.render(ctx, scene, dmats, env, job, stats=False)
            grid = prepare_grid(api, ctx, env, rgba, job, frame)
            job.seed = W.RenderJob().seed
        job.flags = of | tflags
        g2 = api.RenderGraph(ctx, scene, dmats, env, job)
        g2.launch(job.seed)
        g2.stats()
        torch.cuda.synchronize()
        r2, s2 = 0, time.perf_counter()
        for i in range(args.other_steps):
            g2.launch(job.seed)
            st2 = g2.stats()
            r2 += int(st2.extension_rays) + int(st2.shadow_rays)
        e2 = time.perf_counter() - s2
        g2.close()
        others[other] = {"Mrays_per_s": round(r2 / e2 / 1e6, 3), "ms_per_job": round(e2 * 1e3 / args.other_steps, 3),
                         "steps": args.other_steps, "lighting": LIGHTING[other]}
        log(f"{other}: {others[other]}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads()
        log(f"cpu baseline on {threads} threads ({cpu_model()}) ...")
        og = None
        if lflags:
            from oracle import orc_py as orc
            acc, fn, st_ = grid.read_arrays()
            og = orc.LightGrid(cam=job.pos)
            og.accum, og.func = acc.reshape(-1).copy(), fn.reshape(-1).copy()
            og.sat = orc.sat_tiles(og.func, 1600, 1600, 16, 32)
        job.flags = lflags
        cpu = cpu_baseline(W, scene_arrays, mats, rgba, job, npaths, args.cpu_seconds, threads, og,
                           2 if args.lighting == "guided-tile-pdf" else (1 if lflags else 0))
        cpu["cpu_model"] = cpu_model()

    lat_sorted = sorted(lat)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators: scene seed 2, materials seed 2, analytic env map)",
        "config": {
            "workload": "c1 wavefront path tracer: 256 UV spheres x 4096 tris + ground (1,048,578 tris), "
                        f"{job.width}x{job.height} x {job.spp} spp, <=8 bounces, RR from bounce 3, "
                        f"{LIGHTING[args.lighting]}, surviving paths compacted every bounce (appended by the "
                        "shading stage to segmented queues)"
                        + (f"; bounces {args.tail}-{job.max_bounces - 1} in one launch, a path per lane "
                           "(MT_RENDER_TAIL_FROM)" if tflags else ""),
            "tail_from_bounce": args.tail if tflags else None,
            "global_batch": npaths,
            "lighting": args.lighting,
            "traversal_layout": "packed" if job.mode == 0 else "reference",
            "parallelism": "single GPU",
        },
        "samples_per_s": round(npaths * args.steps / elapsed, 1),
        "bounce_paths": bounce_paths,
        "job_latency_ms": qps if qps is not None else {
            "p50": round(lat_sorted[len(lat) // 2], 3),
            "p99": round(lat_sorted[min(len(lat) - 1, int(0.99 * len(lat)))], 3),
            "qps": "sequential (one job in flight)"},
        "step_latency_ms": {"p50": round(lat_sorted[len(lat) // 2], 3),
                            "p99": round(lat_sorted[min(len(lat) - 1, int(0.99 * len(lat)))], 3)},
        "timed_path": "captured render (mt_render_graph) replayed once per step"
                      + ("; the light grid as left by the reference's adaptive phase (frames 0-4 each followed "
                         "by an update), timed frames accumulate into it as later frames do" if lflags else ""),
        "uncaptured": {"Mrays_per_s": round(e_rays / e_elapsed / 1e6, 3),
                       "ms_per_job": round(e_elapsed * 1e3 / args.steps, 3),
                       "note": "mt_render per step (state in the context's reused workspace), HIP events on "
                               "every traversal launch; source of the roofline launch duration"},
        "other_lighting": others,
        "roofline": {
            "kernel": ("closest-hit traversal, packed layout: k_trace_camera (bounce 0, camera rays made in "
                       f"place) and k_trace_pool (bounces 1-{(args.tail if tflags else job.max_bounces) - 1}, "
                       "wave pools with lane refill), averaged over the job's launches"
                       + (" (later bounces run inside k_tail)" if tflags else ""))
                      if job.mode == 0 else "closest-hit traversal (reference layout)",
            "bound": "hbm",
            "achieved": round(achieved, 1) if achieved is not None else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved is not None else None,
            "traffic": traffic,
            "achieved_basis": "measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "gather-calibrated, profiles/) / live average launch duration",
            "alg_achieved": round(alg_achieved, 1),
            "alg_frac": round(alg_achieved / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": round(per_launch_bytes),
            "alg_basis": "reference-layout bytes touched, 40 + 148 I + 116 L + 28 P per ray, no cache credit "
                         "(upper BVH levels are L2/MALL-resident, so this exceeds HBM traffic)",
            "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "valu_issue": valu_issue(valu_from_profiles("c1") if job.mode == 0 else None, avg_launch_s),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def valu_issue(valu, launch_s, source: str = "profiles/counters_c1.md (SQ_INSTS_VALU, GRBM_GUI_ACTIVE)"):
    """The closest-hit launches' VALU issue rate (profiled instructions per launch over the
    live launch time) against the peak wave64 issue rate of the 2-cycle instruction class, with
    the measured cycles per instruction beside it."""
    if valu is None:
        return None
    inst, peak, basis, cpi = valu
    return {"inst_per_launch": round(inst), "achieved_inst_per_s": round(inst / launch_s / 1e9, 1),
            "peak_inst_per_s": round(peak / 1e9, 1), "unit": "G VALU instructions/s",
            "frac": round(inst / launch_s / peak, 4), "peak_basis": basis,
            "cycles_per_inst": cpi,
            "source": source + ", profiles/valu_calibration.md"}


def bench_c2(args) -> int:
    """c2 stress (one step): closest hit of 64*2^20 rays with origins uniform in
    the scene box and directions uniform on the sphere, through 3.6M tiny
    triangles clustered at the centre plus 0.4M large ones spread over a
    200-unit box; then a shadow ray per ray toward one light
    (tMax = dist - 2*EPS_RAY), any hit over them, and the stable compaction
    of the hit rays (ray o/d payloads + source index).  All buffers are
    allocated before the timed region; HIP events on the launch stream time
    each stage.  roofline: closest-hit kernel, algorithmic bytes
    40 + 148*I + 116*L + 28*P per ray counted over all rays by the counting
    traversal (untimed)."""
    import ctypes as C

    import numpy as np
    import torch

    from mambatracer_amd import _lib, api
    from mambatracer_amd import workloads as W

    t0 = time.perf_counter()
    points, faces, nodes, bounds, depth = W.c2_scene()
    n = args.rays
    o, d = api.gen_random_rays(3, 0, n, bounds[0, :3], bounds[0, 4:7], nthreads=16)
    light = np.ascontiguousarray(W.c2_light(bounds), dtype=np.float32)
    cpu = None
    if not args.no_cpu_baseline:   # before the timed region, on the same rays
        threads = cpu_threads()
        cpu = cpu_baseline_c2(points, faces, nodes, bounds, o, d, light, args.cpu_seconds, threads)
        cpu["cpu_model"] = cpu_model()
        log(f"cpu baseline: {cpu}")
    log(f"c2: {faces.shape[0]} triangles, BVH depth {depth}, {n} rays, host prep {time.perf_counter() - t0:.1f} s")
    ctx = api.Context(0)
    scene = api.Scene(ctx, points, faces, nodes, bounds, depth)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    del o, d
    dev = ctx.torch_device
    length = ctx.scalar_i32(n)
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    tt = torch.empty(n, dtype=torch.float32, device=dev)
    occ = torch.empty(n, dtype=torch.int32, device=dev)
    so, sd = torch.empty_like(do), torch.empty_like(dd)
    valid = torch.empty(n, dtype=torch.int32, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    cnt = ctx.scalar_i32(0)
    co, cd = torch.empty_like(do), torch.empty_like(dd)
    pay = (_lib.Payload * 2)()
    pay[0] = _lib.Payload(do.data_ptr(), co.data_ptr(), 16)
    pay[1] = _lib.Payload(dd.data_ptr(), cd.data_ptr(), 16)
    P = lambda x: C.c_void_p(x.data_ptr())
    h = ctx.handle

    modes = {"closest": 0, "any": 0}

    def closest(st=None):
        api.call("mt_trace_closest", h, scene.handle, P(do), P(dd), P(length), n, P(prim), P(tt),
                 P(st) if st is not None else None, 0 if st is not None else modes["closest"])

    def shadow():
        api.call("mt_shadow_rays", h, P(do), P(dd), P(prim), P(tt), light.ctypes.data_as(C.c_void_p), P(length), n,
                 P(so), P(sd), P(valid))

    def anyhit(st=None):
        api.call("mt_trace_any", h, scene.handle, P(so), P(sd), P(length), n, P(occ),
                 P(st) if st is not None else None, 0 if st is not None else modes["any"])

    def compact():
        api.call("mt_compact", h, P(prim), 1, P(length), n, P(idx), P(cnt), pay, 2)

    # untimed: traversal counters over all rays (algorithmic bytes)
    st = torch.empty((n, 3), dtype=torch.int32, device=dev)
    closest(st)
    ilp_c = st.sum(0, dtype=torch.int64).tolist()
    shadow()
    anyhit(st)
    ilp_a = st.sum(0, dtype=torch.int64).tolist()
    del st
    alg_c = 40 * n + 148 * ilp_c[0] + 116 * ilp_c[1] + 28 * ilp_c[2]
    alg_a = 36 * n + 148 * ilp_a[0] + 116 * ilp_a[1] + 28 * ilp_a[2]
    log(f"closest: I {ilp_c[0] / n:.1f} L {ilp_c[1] / n:.1f} P {ilp_c[2] / n:.1f} per ray ({alg_c / n:.0f} B/ray); "
        f"any: I {ilp_a[0] / n:.1f} L {ilp_a[1] / n:.1f} P {ilp_a[2] / n:.1f} per ray")

    # untimed: the traversal mode of each stage, picked by the autotuner on
    # this batch (every mode returns the same bits)
    for key, any_hit, (o_, d_) in (("closest", False, (do, dd)), ("any", True, (so, sd))):
        m, ms = C.c_int32(0), C.c_float(0.0)
        api.call("mt_trace_autotune", h, scene.handle, P(o_), P(d_), P(length), n, 1 if any_hit else 0,
                 P(prim) if not any_hit else None, P(tt) if not any_hit else None, P(occ) if any_hit else None,
                 2, C.byref(m), C.byref(ms))
        modes[key] = int(m.value)
    log(f"autotuned traversal modes: {modes}")
    closest()   # restore this batch's closest hits (the autotuner's last run may have been any hit)
    shadow()

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    stage = np.zeros(4)

    def step():
        ev[0].record()
        closest()
        ev[1].record()
        shadow()
        ev[2].record()
        anyhit()
        ev[3].record()
        compact()
        ev[4].record()

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        stage += [ev[k].elapsed_time(ev[k + 1]) for k in range(4)]
    elapsed = time.perf_counter() - t_start
    ctx.check_scans()
    hits = int((prim >= 0).sum().item())
    survivors = int(cnt.item())
    if survivors != hits:
        log(f"compaction count {survivors} != hits {hits}")
        return 1
    stage /= args.steps
    value = 2 * n * args.steps / elapsed / 1e9
    alg_achieved = alg_c / (stage[0] / 1e3) / 1e9
    log(f"per step: closest {stage[0]:.2f} ms, shadow setup {stage[1]:.2f} ms, any {stage[2]:.2f} ms, "
        f"compaction {stage[3]:.3f} ms; hits {hits}")
    cbytes = n * (4 + 32) + survivors * (32 + 4)
    traffic = traffic_from_profiles("c2", modes["closest"])
    out = {
        "metric": "Grays/s, c2 stress (closest hit + any hit of 64M incoherent rays, 4M triangles)",
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators: scene seed 3, rays seed 3)",
        "config": {"workload": f"c2: {faces.shape[0]} triangles (3.6M tiny in a 0.1 cube + 0.4M large in a "
                               f"200-unit box), {n} rays uniform in the scene box, uniform directions; closest hit, "
                               "a shadow ray per ray to one light, any hit, stable compaction of the hit rays",
                   "rays": n, "traversal_layout": "packed",
                   "traversal_modes": {k: {3: "short stack", 4: "persistent waves with per-lane refill",
                                           6: "wave pools with lane refill",
                                           7: "rays counting-sorted by origin cell (32^3 Morton), then "
                                              "wave pools over that order"}.get(v, v)
                                       for k, v in modes.items()}},
        "stages_ms": {"closest": round(stage[0], 3), "shadow_setup": round(stage[1], 3),
                      "any": round(stage[2], 3), "compaction": round(stage[3], 4)},
        "closest_Grays_per_s": round(n / stage[0] / 1e6, 4),
        "any_Grays_per_s": round(n / stage[2] / 1e6, 4),
        "compaction_GB_per_s": round(cbytes / stage[3] / 1e6, 1),
        "roofline": {
            "kernel": "closest-hit traversal (packed layout, "
                      + {4: "k_trace_refill", 6: "k_trace_pool", 7: "k_bin_hist + k_bin_scatter + k_trace_pool"}
                      .get(modes["closest"], "k_trace") + ")",
            "bound": "hbm",
            "achieved": round(traffic / (stage[0] / 1e3) / 1e9, 1) if traffic else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(traffic / (stage[0] / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "traffic": traffic,
            "achieved_basis": "measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "gather-calibrated, profiles/) / live launch duration",
            "alg_achieved": round(alg_achieved, 1),
            "alg_frac": round(alg_achieved / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": int(alg_c),
            "alg_basis": "reference-layout bytes touched, 40 + 148 I + 116 L + 28 P per ray, no cache credit",
            "avg_launch_ms": round(stage[0], 4),
            "valu_issue": valu_issue(valu_from_profiles("c2", modes["closest"]), stage[0] / 1e3,
                                     "profiles/traffic_trace_c2.md (SQ_INSTS_VALU, GRBM_GUI_ACTIVE)"),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_compact(args) -> int:
    """Stable compaction (mt_compact) alone, outputs preallocated, on the
    shapes the paths use: the 33M-slot render queue after bounce 0 (45%
    survivors; 4 + 16 + 16 B payloads), the c2 hit rays (16M, 87%; 2 x 16 B
    + index) and a multi-job queue (8M, 60%; 8 + 3 x 16 B).  Each case is
    checked against torch's nonzero + gather and timed with HIP events.
    Bytes: flags + payloads read for every slot, payloads (+ index) written
    for the survivors; a plain device copy of 1 GiB is timed beside them."""
    import ctypes as C

    import torch

    from mambatracer_amd import _lib, api

    ctx = api.Context(0)
    dev = ctx.torch_device
    g = torch.Generator(device=dev)
    g.manual_seed(7)

    def ev_time(f, reps=10):
        best = 1e9
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            f()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    copy_gbs = 2 * a.numel() * 4 / ev_time(lambda: b.copy_(a)) / 1e6
    del a, b
    cases = {}
    ok_all = True
    for name, n, p, sizes, index in [("render_queue_33M", 33177600, 0.45, (4, 16, 16), False),
                                     ("c2_hits_16M", 1 << 24, 0.87, (16, 16), True),
                                     ("batch_queue_8M", 8 << 20, 0.6, (8, 16, 16, 16), False)]:
        flags = (torch.rand(n, generator=g, device=dev) < p).to(torch.int32)
        srcs = [torch.empty((n, sz // 4), dtype=torch.int32, device=dev).random_(generator=g) for sz in sizes]
        dsts = [torch.empty_like(x) for x in srcs]
        idx = torch.empty(n, dtype=torch.int32, device=dev) if index else None
        length, cnt = ctx.scalar_i32(n), ctx.scalar_i32(0)
        arr = (_lib.Payload * len(srcs))()
        for j, (x, y) in enumerate(zip(srcs, dsts)):
            arr[j] = _lib.Payload(x.data_ptr(), y.data_ptr(), x.shape[1] * 4)
        ip = C.c_void_p(idx.data_ptr()) if index else None
        f = lambda: api.call("mt_compact", ctx.handle, C.c_void_p(flags.data_ptr()), 0, C.c_void_p(length.data_ptr()),
                             n, ip, C.c_void_p(cnt.data_ptr()), arr, len(srcs))
        f()
        torch.cuda.synchronize()
        k = int(cnt.item())
        sel = flags.nonzero().squeeze(1)
        ok = k == sel.numel() and all(torch.equal(y[:k], x[sel]) for x, y in zip(srcs, dsts))
        if index:
            ok = ok and torch.equal(idx[:k], sel.to(torch.int32))
        ok_all &= ok
        ms = ev_time(f)
        psum = sum(sizes)
        byts = n * (4 + psum) + k * (psum + (4 if index else 0))
        cases[name] = {"ms": round(ms, 4), "GB_per_s": round(byts / ms / 1e6, 1), "survivors": k, "n": n,
                       "payload_bytes": list(sizes), "index": index, "correct": ok}
        log(f"{name}: {cases[name]}")
        del flags, srcs, dsts, idx
    # CCompact.execute on the reference's own Intersection records, in place
    # (mt_compact_isects): 1080p path queue (CDeviceGI's global size w*h) and
    # the 16-spp queue.  Bytes: records read once and the survivors written
    # to the temp buffer, then read back; every record of the length written
    # in place plus the pixel ids.
    for name, n, p in [("ccompact_records_2M", 1920 * 1080, 0.55), ("ccompact_records_33M", 33177600, 0.55)]:
        rec = torch.empty((n, 24), dtype=torch.int32, device=dev).random_(generator=g)
        rec[:, 20] = (torch.rand(n, generator=g, device=dev) < p).to(torch.int32)
        rec[:, 21:] = 0
        pix = torch.randperm(n, generator=g, device=dev).to(torch.int32)
        ref_rec, ref_pix = rec.clone(), pix.clone()
        tmp = torch.empty_like(rec)
        length, cnt = ctx.scalar_i32(n), ctx.scalar_i32(0)
        f = lambda: api.call("mt_compact_isects", ctx.handle, C.c_void_p(rec.data_ptr()), C.c_void_p(pix.data_ptr()),
                             C.c_void_p(tmp.data_ptr()), C.c_void_p(length.data_ptr()), n,
                             C.c_void_p(cnt.data_ptr()))
        f()
        torch.cuda.synchronize()
        k = int(cnt.item())
        sel = ref_rec[:, 20].nonzero().squeeze(1)
        ok = k == sel.numel() and torch.equal(rec[:k], ref_rec[sel]) and torch.equal(pix[:k], ref_pix[sel]) \
            and bool((rec[k:, 20] == 0).all()) and bool((pix[k:] == 0).all())
        ok_all &= ok
        del ref_rec, ref_pix, sel
        ms = ev_time(f)    # later runs compact already-compacted buffers: same k, same traffic
        byts = n * 96 + k * (4 + 96 + 96) + n * (96 + 4)
        cases[name] = {"ms": round(ms, 4), "GB_per_s": round(byts / ms / 1e6, 1), "survivors": k, "n": n,
                       "payload_bytes": [96, 4], "index": False, "correct": ok,
                       "note": "in place on 96-byte Intersection records + pixel ids (two kernels)"}
        log(f"{name}: {cases[name]}")
        del rec, pix, tmp
    ctx.check_scans()
    best = max(c["GB_per_s"] for c in cases.values())
    out = {"metric": "stable compaction bandwidth (flags + payload reads, survivor writes)", "value": best,
           "unit": "GB/s", "n_gpus": 1, "higher_is_better": True, "dtype": "int32/bytes", "data": "synthetic",
           "config": {"workload": "mt_compact on three queue shapes, outputs preallocated; mt_compact_isects "
                                  "in place on Intersection records"},
           "cases": cases, "copy_1GiB_GB_per_s": round(copy_gbs, 1),
           "roofline": {"bound": "hbm", "achieved": best, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(best / HBM_PEAK_GBS, 4)}}
    print(json.dumps(out), flush=True)
    return 0 if ok_all else 1


def bench_stages(args) -> int:
    """Stages B and C alone on full-size inputs, each timed with HIP events
    (best of 10) against its bytes moved: BSDF setup / sample / eval over
    16M hit points (the 1080p x 16 spp bounce-0 hit count), the light-grid
    update and SAT build over the 1600^2 grid, full-map and light-grid
    guided environment sampling and radiance over 16M directions.  Bytes
    count each kernel's own reads and writes (material table, SATs and
    sampling tables are cache-resident and not counted).  Outputs are
    checked for finiteness and the expected pdf sign; parity is the tests'."""
    import numpy as np
    import torch

    from mambatracer_amd import api
    from mambatracer_amd import workloads as W

    ctx = api.Context(0)
    dev = ctx.torch_device
    g = torch.Generator(device=dev)
    g.manual_seed(9)

    def ev_time(f, reps=10):
        best = 1e9
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            f()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    def unit(n):
        v = torch.randn((n, 4), generator=g, device=dev)
        v[:, 3] = 0
        return (v / v[:, :3].norm(dim=1, keepdim=True)).contiguous()

    cases = {}
    ok_all = True

    def add(name, ms, nbytes, n, ok):
        nonlocal ok_all
        ok_all &= bool(ok)
        cases[name] = {"ms": round(ms, 4), "n": n, "bytes": int(nbytes), "GB_per_s": round(nbytes / ms / 1e6, 1),
                       "frac_of_hbm_peak": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3), "ok": bool(ok)}
        log(f"{name}: {cases[name]}")

    # ---- B: BSDF on 16M hit points
    n = 16 << 20
    mats = ctx.to_device(W.c1_materials())
    nmat = int(W.c1_materials().shape[0])
    d = unit(n)
    nrm = unit(n)
    nrm = torch.where(((nrm[:, :3] * d[:, :3]).sum(1, keepdim=True) > 0), -nrm, nrm).contiguous()
    mat = torch.randint(0, nmat, (n,), generator=g, device=dev, dtype=torch.int32)
    rec, mid = api.bsdf_setup(ctx, d, nrm, nrm, mat, mats)
    add("bsdf_setup", ev_time(lambda: api.bsdf_setup(ctx, d, nrm, nrm, mat, mats)), n * (48 + 4 + 80 + 4), n,
        bool(torch.isfinite(rec).all()))
    u = torch.rand((n, 3), generator=g, device=dev)
    f, nd, pdf, cs = api.bsdf_sample(ctx, rec, mid, mats, u)
    add("bsdf_sample", ev_time(lambda: api.bsdf_sample(ctx, rec, mid, mats, u)), n * (80 + 4 + 12 + 32), n,
        bool((pdf >= 0).all()) and float((pdf > 0).float().mean()) > 0.5)
    wi = nd.contiguous()
    add("bsdf_eval", ev_time(lambda: api.bsdf_eval(ctx, rec, mid, mats, wi)), n * (80 + 4 + 12 + 20), n,
        bool(torch.isfinite(api.bsdf_eval(ctx, rec, mid, mats, wi)[1]).all()))
    del d, nrm, mat, rec, mid, u, f, nd, pdf, cs, wi

    # ---- C: light grid (1600 x 1600 = 2.56M cells, 5000 cells of 16 x 32 tiles)
    grid = api.CAdaptiveEnvironment(ctx)
    grid.write_accum((np.random.default_rng(9).random((1600, 1600)) ** 3 + 0.01).astype(np.float32))
    cells = 1600 * 1600
    add("lightgrid_update", ev_time(grid.update), cells * 12, cells, True)
    add("lightgrid_build_sat", ev_time(grid.buildSAT), cells * 8, cells, True)
    cell = torch.randint(0, 5000, (n,), generator=g, device=dev, dtype=torch.int32)
    u2 = torch.rand((n, 2), generator=g, device=dev)
    txy, uv, lpdf = grid.sample(cell, u2)
    add("lightgrid_sample", ev_time(lambda: grid.sample(cell, u2)), n * (4 + 8 + 8 + 8 + 4), n,
        bool((lpdf > 0).all()))
    del cell, u2, txy, uv, lpdf

    # ---- C: environment light, 1024 x 512 map
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    u2 = torch.rand((n, 2), generator=g, device=dev)
    dirs, c, pw, tx = env.illuminate(u2)
    add("env_illuminate", ev_time(lambda: env.illuminate(u2)), n * (8 + 16 + 16 + 4 + 4), n,
        bool(torch.isfinite(dirs).all()) and bool((pw >= 0).all()))
    dd = unit(n)
    add("env_radiance", ev_time(lambda: env.radiance(dd)), n * (16 + 16 + 4 + 4), n, True)
    grid.setCameraPosition(W.RenderJob().pos)
    hit = (torch.rand((n, 4), generator=g, device=dev) * 10 - 5).contiguous()
    hit[:, 3] = 0
    u4 = torch.rand((n, 4), generator=g, device=dev)
    _, _, apw, _, agi = env.illuminate_adaptive(grid, hit, u4)
    add("env_illuminate_adaptive", ev_time(lambda: env.illuminate_adaptive(grid, hit, u4)),
        n * (16 + 16 + 16 + 16 + 4 + 4 + 4), n, bool((apw >= 0).all()) and bool((agi >= 0).all()))
    grid.clearBuffers()
    add("env_radiance_adaptive", ev_time(lambda: env.radiance_adaptive(grid, dd, hit)),
        n * (16 + 16 + 16 + 4 + 4 + 4 + 4), n, True)
    out = {"metric": "stage B/C kernels on full-size inputs (bytes moved per second)",
           "value": cases["bsdf_sample"]["GB_per_s"], "unit": "GB/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f32", "data": "synthetic (seeded)",
           "config": {"workload": "BSDF setup/sample/eval on 16M hits; light-grid update/SAT/sampling; env "
                                  "sampling and radiance (full map and light-grid guided) on 16M directions"},
           "cases": cases}
    print(json.dumps(out), flush=True)
    return 0 if ok_all else 1


def bench_bvh(args) -> int:
    """The device BVH builder (mt_bvh_build_device, CNormalBVH.build's
    midpoint split) on the c1 (1,048,578 triangles) and c2 (4,000,000)
    scenes from device points/faces, against the host builder; the device
    nodes must equal the host builder's.  Best of 3 builds each."""
    import numpy as np
    import torch

    from mambatracer_amd import api

    ctx = api.Context(0)
    res = {}
    ok_all = True
    for name in ("c1", "c2"):
        if name == "c1":
            points, _, faces = api.gen_spheres(2, 256)
        else:
            tp, tf = api.gen_soup(3, 3_600_000, -0.05, 0.05, 1e-4, 1e-3, first_vertex=0)
            lp, lf = api.gen_soup(1003, 400_000, -100.0, 100.0, 1.0, 10.0, first_vertex=tp.shape[0])
            points, faces = np.concatenate([tp, lp]), np.concatenate([tf, lf])
        t0 = time.perf_counter()
        hn, hb, hd = api.CNormalBVH.build(points, faces)
        host_ms = (time.perf_counter() - t0) * 1e3
        dp, df = ctx.to_device(points), ctx.to_device(faces)
        api.CNormalBVH.build_device(ctx, dp, df)
        torch.cuda.synchronize()
        best = 1e9
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dn, db, dd = api.CNormalBVH.build_device(ctx, dp, df)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        same = bool(np.array_equal(dn.cpu().numpy(), hn)) and dd == hd
        ok_all &= same
        res[name] = {"triangles": int(faces.shape[0]), "device_ms": round(best * 1e3, 2),
                     "host_ms_multithreaded": round(host_ms, 1), "depth": dd, "same_nodes_as_host": same}
        log(f"{name}: {res[name]}")
    out = {"metric": "device BVH build time (midpoint split, reference node order)", "value": res["c2"]["device_ms"],
           "unit": "ms (4M triangles)", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
           "data": "synthetic (c1 and c2 scene generators)", "config": {"workload": "mt_bvh_build_device on c1, c2"},
           "builds": res}
    print(json.dumps(out), flush=True)
    return 0 if ok_all else 1


def bench_jobs(args) -> int:
    """c3 (all jobs at t=0) or c4 (Poisson arrivals): the multi-job engine
    against one-job-at-a-time plain renders of the same requests, images
    compared bit for bit."""
    import numpy as np
    import torch

    from mambatracer_amd import api, serve
    from mambatracer_amd import workloads as W

    t0 = time.perf_counter()
    ctx = api.Context(0)
    dmats = ctx.to_device(W.c1_materials())
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    reqs = []
    if args.workload == "c3":
        specs, ntris = W.c3_specs(n_jobs=args.jobs)
        scenes = []
        for spec, n in zip(specs, ntris):
            scenes.append(api.Scene(ctx, *W.soup_scene(n, spec.seed)))
            reqs.append(serve.Request(0.0, scenes[-1], dmats, env, W.soup_job(spec), tag=spec.index))
        desc = (f"c3: {len(specs)} jobs submitted at t=0, side log-uniform 64-1024, spp 1-64, each on its own "
                f"soup of 1k-500k triangles ({sum(ntris)} total), c1 materials and sky, <=8 bounces")
    else:
        sizes = W.c4_soup_sizes()
        scenes = [api.Scene(ctx, *W.soup_scene(n, 5_000 + i)) for i, n in enumerate(sizes)]
        points, normals, faces, nodes, bounds, depth = W.c1_scene()
        scenes.append(api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals))
        for t, spec in W.c4_arrivals(args.rate, args.duration, n_soups=len(sizes)):
            job = W.soup_job(spec) if spec.side > 0 else W.RenderJob(seed=spec.seed)
            reqs.append(serve.Request(t, scenes[spec.scene], dmats, env, job, tag=spec.index))
        desc = (f"c4: Poisson arrivals at {args.rate} jobs/s for {args.duration} s ({len(reqs)} jobs); 90% c3-sized "
                f"jobs on {len(sizes)} resident soups, 10% c1 1920x1080x16spp jobs on the 1M-triangle scene")
    torch.cuda.synchronize()
    paths = sum(r.job.width * r.job.height * r.job.spp for r in reqs)
    log(f"{len(reqs)} jobs, {paths} paths, setup {time.perf_counter() - t0:.1f} s")

    # warm both paths (code objects, allocator) on two small jobs
    warm = [serve.Request(0.0, reqs[k].scene, dmats, env, W.soup_job(W.JobSpec(k, 0, 32, 1, 7)), tag=k)
            for k in range(2)]
    serve.serve_sequential(ctx, warm)
    serve.serve_batched(ctx, warm, queue_capacity=args.queue)

    # the engine's settings for this workload, picked by the autotuner on a
    # sample of it (untimed): c3 every 4th job for jobs/s, c4 the first 1.5 s
    # of arrivals for p99 latency
    if args.workload == "c3":
        tsample, tobj = reqs[::4], "jobs_per_s"
    else:
        tsample, tobj = [r for r in reqs if r.arrival < 1.5], "p99"
    tuned, trials = serve.autotune_batch(ctx, tsample, objective=tobj)
    log(f"autotuned engine settings: {tuned} from {len(trials)} trials")

    seq, seq_rays = serve.serve_sequential(ctx, reqs, keep_images=True)
    s_seq = serve.summary(seq, seq_rays)
    log(f"sequential: {s_seq}")
    bt = api.Batch(ctx, args.queue, 32 << 30, 1024, policy=args.policy)
    bt.set_iterations(args.iters)
    bat, st = serve.serve_batched(ctx, reqs, queue_capacity=args.queue, keep_images=True, batch=bt)
    st = bt.stats()
    bt.close()
    s_bat = serve.summary(bat, int(st.extension_rays) + int(st.shadow_rays))
    log(f"batched: {s_bat}, iterations {int(st.iterations)}, arena peak {int(st.arena_peak) / 1e9:.2f} GB")
    ref = {d.tag: d.image for d in seq}
    mismatched = sum(0 if torch.equal(ref[d.tag], d.image) else 1 for d in bat)
    modes = {"batched": s_bat, "sequential": s_seq}
    sjf, sjf_rays = serve.serve_sequential(ctx, reqs, keep_images=True, order="sjf")
    modes["sequential_sjf"] = serve.summary(sjf, sjf_rays)
    log(f"sequential, shortest job first: {modes['sequential_sjf']}")
    mismatched += sum(0 if torch.equal(ref[d.tag], d.image) else 1 for d in sjf)
    two, two_rays = serve.serve_two_lane(ctx, reqs, keep_images=True)
    modes["two_lane"] = serve.summary(two, two_rays)
    log(f"two lanes (large jobs / small jobs shortest first): {modes['two_lane']}")
    mismatched += sum(0 if torch.equal(ref[d.tag], d.image) else 1 for d in two)
    bt = serve.make_batch(ctx, tuned)
    batu, _ = serve.serve_batched(ctx, reqs, keep_images=True, batch=bt)
    st_t = bt.stats()
    bt.close()
    modes["batched_autotuned"] = serve.summary(batu, int(st_t.extension_rays) + int(st_t.shadow_rays))
    log(f"batched, autotuned settings: {modes['batched_autotuned']}")
    mismatched += sum(0 if torch.equal(ref[d.tag], d.image) else 1 for d in batu)
    twob, twob_rays = serve.serve_two_lane(ctx, reqs, keep_images=True, small_order="batched",
                                           small_iterations=args.iters)
    modes["two_lane_batched"] = serve.summary(twob, twob_rays)
    log(f"two lanes (large jobs / small jobs continuously batched): {modes['two_lane_batched']}")
    mismatched += sum(0 if torch.equal(ref[d.tag], d.image) else 1 for d in twob)
    # Amdahl bound of the multi-job gain on all-at-once arrivals: a job of at
    # least `fill` paths keeps the GPU busy alone, so no mix shortens its own
    # render; batching can at best run every smaller job at the rate of the
    # largest jobs.  bound = total sequential time / (time of the large jobs
    # alone + paths of the small jobs / the large jobs' path rate).
    svc = {}
    prev = 0.0
    for d in sorted(seq, key=lambda d: d.complete):
        svc[d.tag] = d.complete - max(prev, d.arrival)
        prev = d.complete
    paths_of = {r.tag: r.job.width * r.job.height * r.job.spp for r in reqs}
    fill = 1 << 22
    big = [t for t in svc if paths_of[t] >= fill]
    small = [t for t in svc if paths_of[t] < fill]
    t_big = sum(svc[t] for t in big)
    rate_big = sum(paths_of[t] for t in big) / t_big if t_big > 0 else None
    amdahl = None
    if rate_big:
        t_lower = t_big + sum(paths_of[t] for t in small) / rate_big
        amdahl = {"fill_paths": fill, "large_jobs": len(big), "large_share_of_paths":
                  round(sum(paths_of[t] for t in big) / sum(paths_of.values()), 3),
                  "large_jobs_time_s": round(t_big, 4), "small_jobs_time_s": round(sum(svc[t] for t in small), 4),
                  "large_path_rate": round(rate_big, 1), "bound_speedup": round(sum(svc.values()) / t_lower, 3)}
        log(f"Amdahl bound: {amdahl}")
    ctx.sync()
    metric = "jobs/s, multi-job serving vs one job at a time (best mode; every mode reported)" \
        if args.workload == "c3" else "p50/p99 render-job latency at fixed QPS (best serving mode; every mode reported)"
    if args.workload == "c3":
        best = max(modes, key=lambda m: modes[m]["jobs_per_s"])
    else:
        best = min(modes, key=lambda m: modes[m]["latency_ms"]["p99"])
    out = {
        "metric": metric,
        "value": modes[best]["jobs_per_s"] if args.workload == "c3" else modes[best]["latency_ms"]["p99"],
        "unit": "jobs/s" if args.workload == "c3" else "ms (p99)",
        "value_mode": best,
        "n_gpus": 1,
        "higher_is_better": args.workload == "c3",
        "dtype": "f32",
        "data": "synthetic (seeded generators)",
        "config": {"workload": desc, "queue_capacity": args.queue, "paths": paths,
                   "policy": "smallest job first" if args.policy == 1 else "oldest job first",
                   "iterations_per_step": args.iters},
        "amdahl": amdahl,
        "autotune": {"choice": tuned, "objective": tobj,
                     "trials": [{"setting": t, tobj: r} for t, r in trials],
                     "sample": "every 4th request, submitted at once" if tobj == "jobs_per_s"
                     else "the requests arriving in the first 1.5 s, at their arrival times"},
        **modes,
        "speedup_jobs_per_s": {m: round(modes[m]["jobs_per_s"] / s_seq["jobs_per_s"], 3) for m in modes},
        "images_mismatched": mismatched,
        "images_compared": len(bat) * (len(modes) - 1),
    }
    print(json.dumps(out), flush=True)
    return 0 if mismatched == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
