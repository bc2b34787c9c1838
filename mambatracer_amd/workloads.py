"""Seeded synthetic workloads: the benchmark configurations of the hot path.

  c0  room box [-1,1]^3 (12 tris) + 20 000 random triangles (centroids in
      [-0.8,0.8]^3, edges 0.01-0.05), seed 1; 512x512 pinhole camera at
      (0,0,-0.95) looking +z, 60 deg vertical FOV, 1 spp at pixel centres.
  c1  256 UV spheres x 4096 triangles (radii lognormal(-2.5, 0.5), centres in
      [-5,5]^3) + a 2-triangle ground quad, seed 2; GGX materials (roughness
      U[0.05,1], albedo U[0.2,0.9]^3, F0 0.04); 1024x512 float32 equirect
      sky gradient 0.2-1.0 plus a Gaussian sun (sigma 1 deg, peak 1e4);
      1920x1080, 16 spp, up to 8 bounces, Russian roulette after bounce 3.
  c2  4M-triangle soup: 3.6M tiny (edges 1e-4-1e-3) in a 0.1-wide cube,
      0.4M large (edges 1-10) in [-100,100]^3, seed 3; rays with origins
      uniform in the scene box and directions uniform on the sphere; a point
      light uniform in the scene box.
The generators are the library's own (mt_gen_*), deterministic per seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import api

MAT_WORDS = 36


def c0_scene():
    rp, rf = api.gen_room(0)
    sp, sf = api.gen_soup(1, 20000, -0.8, 0.8, 0.01, 0.05, first_vertex=8, mat=0)
    points = np.concatenate([rp, sp])
    faces = np.concatenate([rf, sf])
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    return points, faces, nodes, bounds, depth


def c0_rays(w: int = 512, h: int = 512):
    return api.gen_camera_rays([0, 0, -0.95], [0, 0, 1], [0, 1, 0], 60.0, w, h)


def material_records(diffuse, glossy_f0, alpha_x, alpha_y, emission=None, emission_power=None):
    """Packs surface parameters into 36-word records (Bsdf.cl:24-42 layout)."""
    n = len(diffuse)
    m = np.zeros((n, MAT_WORDS), np.float32)
    m[:, 4:7] = diffuse
    m[:, 8] = 1.0                     # diffuse_param.x = scale
    m[:, 12:15] = glossy_f0
    m[:, 16] = 1.0                    # glossy_param.x = scale
    m[:, 17] = alpha_x                # glossy_param.y = ax
    m[:, 18] = alpha_y                # glossy_param.z = ay
    m[:, 20:23] = 0.95                # mirror_color (unused by the two lobes)
    m[:, 28:31] = 1.0                 # emission_color, CSurfaceParameter2 default
    if emission is not None:
        m[:, 28:31] = emission
        m[:, 32] = 1.0                # emission_param.x = scale
        m[:, 33] = emission_power     # emission_param.y = power
    return m


def c1_materials(count: int = 256, seed: int = 2):
    rng = np.random.default_rng(seed)
    albedo = rng.uniform(0.2, 0.9, size=(count + 1, 3))
    rough = rng.uniform(0.05, 1.0, size=count + 1)
    albedo[count] = (0.5, 0.5, 0.5)   # ground
    rough[count] = 0.5
    return material_records(albedo, 0.04, rough, rough)


def env_map(w: int = 1024, h: int = 512, sun_u: float = 0.3, sun_v: float = 0.28, sigma_deg: float = 1.0,
            peak: float = 1e4):
    """Equirect map in the reference's parameterisation (getSphericalDirection,
    Util.cl:423-439): sky gradient 0.2 (nadir) .. 1.0 (zenith) + Gaussian sun."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    u, v = (x + 0.5) / w, (y + 0.5) / h
    phi, theta = math.pi * v, 2 * math.pi * u
    d = np.stack([-np.sin(phi) * np.cos(theta), np.cos(phi), np.sin(phi) * np.sin(theta)], -1)
    sp, st = math.pi * sun_v, 2 * math.pi * sun_u
    s = np.array([-math.sin(sp) * math.cos(st), math.cos(sp), math.sin(sp) * math.sin(st)])
    ang = np.arccos(np.clip(d @ s, -1, 1))
    sig = math.radians(sigma_deg)
    val = 0.2 + 0.8 * (1.0 - v) + peak * np.exp(-0.5 * (ang / sig) ** 2)
    rgba = np.zeros((h, w, 4), np.float32)
    rgba[..., 0] = val
    rgba[..., 1] = val
    rgba[..., 2] = val
    return rgba


@dataclass
class RenderJob:
    width: int = 1920
    height: int = 1080
    spp: int = 16
    max_bounces: int = 8
    rr_start: int = 3
    seed: int = 2
    pos: tuple = (0.0, 2.0, -15.0)
    look: tuple = (0.0, -1.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    vfov_deg: float = 45.0
    mode: int = 0
    flags: int = 0


def c1_scene(count: int = 256, seed: int = 2):
    points, normals, faces = api.gen_spheres(seed, count, -2.5, 0.5, -5.0, 5.0, -5.0, 50.0)
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    return points, normals, faces, nodes, bounds, depth


def c2_scene(n_tiny: int = 3_600_000, n_large: int = 400_000, seed: int = 3):
    tp, tf = api.gen_soup(seed, n_tiny, -0.05, 0.05, 1e-4, 1e-3, first_vertex=0)
    lp, lf = api.gen_soup(seed + 1000, n_large, -100.0, 100.0, 1.0, 10.0, first_vertex=tp.shape[0])
    points = np.concatenate([tp, lp])
    faces = np.concatenate([tf, lf])
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    return points, faces, nodes, bounds, depth


def c2_light(bounds, seed: int = 3):
    rng = np.random.default_rng(seed)
    lo, hi = bounds[0, :3].astype(np.float64), bounds[0, 4:7].astype(np.float64)
    return (lo + (hi - lo) * rng.random(3)).astype(np.float32)


# ---------------------------------------------------------------------------
# Batched jobs (c3) and the serving loop (c4)
# ---------------------------------------------------------------------------
@dataclass
class JobSpec:
    """One render job of the multi-job workloads."""
    index: int
    scene: int          # index into the scene list the job renders
    side: int           # square image side
    spp: int
    seed: int


def _log_uniform(rng, lo: float, hi: float) -> float:
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


def soup_scene(num_tris: int, seed: int, n_materials: int = 257):
    """A c0-style random-triangle soup (centroids in [-0.8,0.8]^3, edges
    0.01-0.05) without the room box, so the environment lights it; material
    ids drawn uniformly from the c1 material table."""
    points, faces = api.gen_soup(seed, num_tris, -0.8, 0.8, 0.01, 0.05, first_vertex=0, mat=0)
    faces[:, 9] = np.random.default_rng(seed).integers(0, n_materials, size=num_tris, dtype=np.int32)
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    return points, faces, nodes, bounds, depth


def c3_specs(n_jobs: int = 256, seed: int = 4, side=(64, 1024), spp=(1, 64), tris=(1_000, 500_000)):
    """c3: jobs of uneven size, each with its own soup: image side log-uniform
    in [side], spp uniform integer in [spp], triangle count log-uniform in
    [tris].  Returns (job specs, per-scene triangle counts)."""
    rng = np.random.default_rng(seed)
    specs, ntris = [], []
    for i in range(n_jobs):
        s = int(round(_log_uniform(rng, side[0], side[1])))
        k = int(rng.integers(spp[0], spp[1] + 1))
        n = int(round(_log_uniform(rng, tris[0], tris[1])))
        specs.append(JobSpec(i, i, s, k, seed * 100_000 + i))
        ntris.append(n)
    return specs, ntris


def soup_job(spec: JobSpec) -> RenderJob:
    """Render job of a soup scene: the c0 camera (pinhole at (0,0,-0.95)
    looking +z, 60 deg), c1 path settings (8 bounces, roulette from 3)."""
    return RenderJob(width=spec.side, height=spec.side, spp=spec.spp, max_bounces=8, rr_start=3,
                     seed=spec.seed, pos=(0.0, 0.0, -0.95), look=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0),
                     vfov_deg=60.0)


def c4_arrivals(rate: float, duration: float, seed: int = 5, n_soups: int = 31, large_frac: float = 0.1,
                side=(64, 1024), spp=(1, 64)):
    """c4: Poisson arrivals at `rate` jobs/s for `duration` s.  A job is a
    c1-scene 1080p x 16 spp render with probability `large_frac` (scene index
    n_soups), else a c3-sized job on one of `n_soups` resident soups.
    Returns a list of (arrival time s, JobSpec)."""
    rng = np.random.default_rng(seed)
    t, out, i = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        if rng.random() < large_frac:
            spec = JobSpec(i, n_soups, -1, 16, seed * 100_000 + i)
        else:
            s = int(round(_log_uniform(rng, side[0], side[1])))
            k = int(rng.integers(spp[0], spp[1] + 1))
            spec = JobSpec(i, int(rng.integers(0, n_soups)), s, k, seed * 100_000 + i)
        out.append((t, spec))
        i += 1


def c4_soup_sizes(n_soups: int = 31, seed: int = 5, tris=(1_000, 500_000)):
    rng = np.random.default_rng(seed + 1)
    return [int(round(_log_uniform(rng, tris[0], tris[1]))) for _ in range(n_soups)]
