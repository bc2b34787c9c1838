"""Closed-form pins of the oracle's bounce loop (orc_render.c), on the CPU.

Furnace: a diffuse plane of albedo rho seen from above under a constant
environment of radiance L, two path segments (camera -> plane, plane ->
sky), no Russian roulette.  Each path's radiance is the sum of the light
sample (DirectLight, MIS weight mis2(p_light, p_bsdf)) and the escaped BSDF
ray (LightHitPass, mis2(p_bsdf, p_env)), so the expected pixel is

    E = (rho L / pi) * int_{upper hemisphere} cos [k(w) w_N(w) + w_B(w)] dw

with k the ratio of the true sampling density to the pdf the estimator
divides by (1 for the full-map light and for the guided light with the tile
pick's pdf; the tile pick's continuous pdf pdfLL for the reference's guided
pdf), w_N = p_N / (p_N + cos/pi) and w_B = (cos/pi) / (cos/pi + p_E), where
the escaped-ray pdf p_E is taken at the top edge of the ray's texel row
(sin(pi y / h), Light.cl:283-284; the light sample's pdf uses the sampled
angle).  E is integrated here independently in float64, and the oracle's
Monte Carlo mean must match it within 4 standard errors.  For the full-map
and tile-pdf lights E is within 1% of rho L on a 128-row map (the row-edge pdf is the only
departure); the reference's guided pdf is biased by the tile-pick density, and
its mean lands on the biased E, far from rho L.

Mesh light: a one-sided triangle emitter of radiance Le above the plane,
no environment; the pixel's expected radiance is rho / pi times the
irradiance given by Lambert's polygon formula.

Glossy pdf: the reference evaluates the VNDF pdf with the generated
direction (pdfGGXVNDF(wo, wh), Bsdf.cl:417,492); multiplied by
G1(wi) wo.z / (G1(wo) wi.z) it must equal the density of the direction
actually sampled, D_wi(h) / (4 wi.h), with D and G1 from the textbook
Trowbridge-Reitz formulas in float64."""
import math

import numpy as np
import pytest

from oracle import orc_py as orc


class Job:
    def __init__(self, **kw):
        self.width, self.height, self.spp = 32, 32, 64
        self.max_bounces, self.rr_start, self.seed = 2, -1, 7
        self.pos, self.look, self.up, self.vfov_deg = (0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 20.0
        self.mode, self.flags = 0, 0
        self.__dict__.update(kw)


def material(diffuse, emission=None, power=0.0):
    m = np.zeros(36, np.float32)
    m[4:7] = diffuse
    m[8] = 1.0
    m[16] = 1.0
    m[17] = m[18] = 0.5
    if emission is not None:
        m[28:31] = emission
        m[32] = 1.0
        m[33] = power
    return m


def plane_scene(extra_tris=(), extra_mats=()):
    """Quad y = 0 over [-100, 100]^2 (material 0) plus extra triangles."""
    pts = [(-100, 0, -100), (100, 0, -100), (100, 0, 100), (-100, 0, 100)]
    faces = [(0, 1, 2), (0, 2, 3)]
    mats_idx = [0, 0]
    for k, tri in enumerate(extra_tris):
        base = len(pts)
        pts.extend(tri)
        faces.append((base, base + 1, base + 2))
        mats_idx.append(1 + k)
    points = np.zeros((len(pts), 4), np.float32)
    points[:, :3] = pts
    F = np.full((len(faces), 10), -1, np.int32)
    F[:, :3] = faces
    F[:, 9] = mats_idx
    nodes, bounds = orc.bvh_build(points, F)[:2]
    mats = np.stack([material(0.5)] + list(extra_mats))
    return points, F, nodes, bounds, mats


def constant_env(w=128, h=64, value=1.0):
    rgba = np.zeros((h * w, 4), np.float32)
    rgba[:, :3] = value
    lum, sat = orc.env_lum_sat(rgba, w, h)
    return (rgba, lum, sat, w, h)


def render_mean(scene, env, job, grid=None, grid_mode=0):
    points, F, nodes, bounds, mats = scene
    n = job.width * job.height * job.spp
    L, _ = orc.render_paths(points, None, F, nodes, bounds, mats, env, job, np.arange(n, dtype=np.int32),
                            grid=grid, grid_mode=grid_mode)
    v = L[:n, 0].astype(np.float64)
    return v.mean(), v.std(ddof=1) / math.sqrt(n)


def furnace_expectation(tile_k, w, h, rho=0.5, Lenv=1.0, corrected=False):
    """E above by quadrature over the upper hemisphere (the plane's normal is
    +y, the map's polar axis): phi = polar angle in (0, pi/2), theta the
    azimuth; tile_k[ty, tx] = pdfLL of each grid tile (1 everywhere for the
    full-map light)."""
    rows = h // 2                                    # texel rows of the upper hemisphere
    sub = 400
    total = 0.0
    for y in range(rows):
        phi = (y + (np.arange(sub) + 0.5) / sub) * math.pi / h
        dphi = math.pi / h / sub
        cos, sin = np.cos(phi), np.sin(phi)
        p_b = cos / math.pi
        ty = int((y / h) * 32)                       # tile row of this texel row
        for tx in range(16):                         # tile columns: equal azimuth spans
            k = tile_k[ty, tx]
            dth = 2 * math.pi / 16
            p_true = k / (2 * math.pi ** 2 * sin)
            p_n = p_true if corrected else 1.0 / (2 * math.pi ** 2 * sin)
            if y == 0:   # sin(0) = 0: the escaped-ray pdf is infinite, its weight 0
                w_b = np.zeros_like(cos)
            else:
                p_e_base = 1.0 / (2 * math.pi ** 2 * math.sin(math.pi * y / h))
                p_e = k * p_e_base if corrected else p_e_base
                w_b = p_b / (p_b + p_e)
            if k > 0:
                nee = (p_true / p_n) * (p_n / (p_n + p_b))
            else:
                nee = np.zeros_like(cos)
            total += np.sum(cos * (nee + w_b) * sin) * dphi * dth
    return rho * Lenv / math.pi * total


def test_furnace_full_map_light():
    scene = plane_scene()
    env = constant_env()
    job = Job()
    mean, se = render_mean(scene, env, job)
    E = furnace_expectation(np.ones((32, 16)), env[3], env[4])
    print(f"full-map furnace: mean {mean:.5f} +- {se:.5f}, E {E:.5f}, rho L 0.5")
    assert abs(E - 0.5) < 0.5 * 1e-2
    assert abs(mean - E) < 4 * se


def _tile_func(seed, zero_frac=0.25):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.2, 3.0, 512).astype(np.float32)
    t[rng.random(512) < zero_frac] = 0.0
    return t


def _tile_k(tile):
    t = tile.astype(np.float64).reshape(32, 16)      # entry i = tile (i % 16, i / 16)
    return t * 512.0 / t.sum()


@pytest.mark.parametrize("corrected", [True, False])
def test_furnace_guided_light(corrected):
    scene = plane_scene()
    env = constant_env()
    if corrected:   # random weights, a quarter of the tiles never picked
        tile = _tile_func(3)
    else:           # the upper hemisphere's tiles three times as likely: pdfLL = 1.5 there
        tile = np.where(np.arange(512) // 16 < 16, 3.0, 1.0).astype(np.float32)
    grid = orc.LightGrid(tile, cam=(0.0, 1.0, 0.0))
    job = Job(seed=11)
    mean, se = render_mean(scene, env, job, grid=grid, grid_mode=2 if corrected else 1)
    E = furnace_expectation(_tile_k(tile), env[3], env[4], corrected=corrected)
    print(f"guided furnace corrected={corrected}: mean {mean:.5f} +- {se:.5f}, E {E:.5f}")
    assert abs(mean - E) < 4 * se
    if corrected:
        assert abs(E - 0.5) < 0.5 * 1e-2
    else:
        assert abs(E - 0.5) > 20 * se        # the reference's pdf is biased, measurably
    grid.fold()
    assert grid.accum.sum() > 0                # escaped rays and light samples accumulated


def lambert_polygon_irradiance(x, n, verts, Le):
    """Irradiance at x (normal n) from a uniform polygon of radiance Le:
    E = Le / 2 * |sum_i angle(v_i, v_i+1) * n . unit(v_i x v_i+1)|."""
    v = [np.asarray(p, np.float64) - x for p in verts]
    s = 0.0
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        ang = math.acos(np.clip(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b), -1, 1))
        c = np.cross(a, b)
        s += ang * np.dot(n, c / np.linalg.norm(c))
    return 0.5 * Le * abs(s)


def test_mesh_light_lambert_irradiance():
    """One emitting triangle facing down at height 1 over the plane; the
    camera sees a ~1e-4-wide spot of the plane at the origin."""
    tri = [(-0.3, 1.0, -0.2), (0.5, 1.0, 0.6), (0.4, 1.0, -0.4)]
    e1 = np.subtract(tri[1], tri[0])
    e2 = np.subtract(tri[2], tri[0])
    if np.cross(e1, e2)[1] > 0:                      # one-sided: normal must face -y
        tri = [tri[0], tri[2], tri[1]]
    Le = 3.0
    scene = plane_scene([tri], [material(0.0, emission=(1.0, 1.0, 1.0), power=Le)])
    light_faces = np.array([2], np.int32)
    job = Job(width=1, height=1, spp=200_000, pos=(0.0, 3.0, -3.0), look=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
              vfov_deg=0.002, seed=5)
    points, F, nodes, bounds, mats = scene
    n = job.spp
    L, _ = orc.render_paths(points, None, F, nodes, bounds, mats, None, job, np.arange(n, dtype=np.int32),
                            light_faces=light_faces)
    v = L[:n, 0].astype(np.float64)
    mean, se = v.mean(), v.std(ddof=1) / math.sqrt(n)
    E = lambert_polygon_irradiance(np.zeros(3), np.array([0.0, 1.0, 0.0]), tri, Le)
    want = 0.5 / math.pi * E
    print(f"mesh light: mean {mean:.6f} +- {se:.6f}, rho/pi * E {want:.6f}")
    assert abs(mean - want) < 4 * se + 1e-4 * want


def ggx_D64(h, ax, ay):
    c2 = h[..., 2] ** 2
    s2 = 1 - c2
    t2 = s2 / c2
    cph = np.where(s2 > 0, h[..., 0] / np.sqrt(np.maximum(s2, 1e-300)), 1.0)
    sph = np.where(s2 > 0, h[..., 1] / np.sqrt(np.maximum(s2, 1e-300)), 0.0)
    e = (cph ** 2 / ax ** 2 + sph ** 2 / ay ** 2) * t2
    return 1.0 / (math.pi * ax * ay * c2 * c2 * (1 + e) ** 2)


def ggx_G1_64(w, ax, ay):
    c2 = w[..., 2] ** 2
    s2 = 1 - c2
    cph = np.where(s2 > 0, w[..., 0] / np.sqrt(np.maximum(s2, 1e-300)), 1.0)
    sph = np.where(s2 > 0, w[..., 1] / np.sqrt(np.maximum(s2, 1e-300)), 0.0)
    a2 = cph ** 2 * ax ** 2 + sph ** 2 * ay ** 2
    lam = (-1 + np.sqrt(1 + a2 * s2 / c2)) / 2
    return 1.0 / (1.0 + lam)


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.6, 0.2)])
def test_glossy_pdf_quirk_relation(ax, ay):
    """EvaluateBsdf on a pure glossy lobe returns the quirk pdf; corrected by
    G1(wi) wo.z / (G1(wo) wi.z) it is the sampled direction's true density."""
    rng = np.random.default_rng(9)
    m = np.zeros((1, 36), np.float32)
    m[0, 12:15] = 0.04
    m[0, 16] = 1.0
    m[0, 17], m[0, 18] = ax, ay
    n = 4000
    # the local frame of a +z normal is the identity (Frisvad, n = z)
    wi = rng.normal(size=(n, 3))
    wi[:, 2] = np.abs(wi[:, 2]) + 0.2
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo[:, 2] = np.abs(wo[:, 2]) + 0.2
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    ray_d = np.zeros((n, 4), np.float32)
    ray_d[:, :3] = -wi                                 # the ray arrives along -wi
    nrm = np.zeros((n, 4), np.float32)
    nrm[:, 2] = 1.0
    rec, mid = orc.bsdf_setup(ray_d, nrm, nrm, np.zeros(n, np.int32), m)
    f, pdf, _ = orc.bsdf_eval(rec, mid, m, wo.astype(np.float32))
    wi64 = np.column_stack([rec[:, 9], rec[:, 10], rec[:, 11]]).astype(np.float64)   # localDirFix as set up
    h = wi64 + wo
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    ok = (wi64[:, 2] > 0.05) & (wo[:, 2] > 0.05) & (np.sum(wi64 * h, 1) > 0.05)
    d_wi = ggx_G1_64(wi64, ax, ay) * np.sum(wi64 * h, 1) * ggx_D64(h, ax, ay) / wi64[:, 2]
    true_pdf = d_wi / (4 * np.sum(wi64 * h, 1))
    fix = ggx_G1_64(wi64, ax, ay) * wo[:, 2] / (ggx_G1_64(wo, ax, ay) * wi64[:, 2])
    got = pdf.astype(np.float64) * fix
    rel = np.abs(got[ok] - true_pdf[ok]) / true_pdf[ok]
    assert ok.sum() > 3000 and rel.max() < 2e-4, rel.max()
