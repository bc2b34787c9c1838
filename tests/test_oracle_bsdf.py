"""Pins the oracle's stage B (GGX/Lambert BSDF) and RNG against textbook
results computed independently in float64 numpy: GGX normalisation, the
visible-normal identity int G1(v) <v,h>+ D(h) dh = v.z (Heitz 2014), the VNDF
sampling density (Heitz 2018), Lambert c/pi, reciprocity of f, and hand-
derived WangHash values with arithmetic shifts."""
import math

import numpy as np
import pytest

from oracle import orc_py as orc


def sph(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def hemisphere_grid(nt=300, nphi=160):
    # theta concentrated near the pole (D is peaked there for small alpha)
    s = (np.arange(nt) + 0.5) / nt
    theta = 0.5 * math.pi * s * s
    dtheta = 0.5 * math.pi * 2 * s / nt
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * math.pi
    T, P = np.meshgrid(theta, phi, indexing="ij")
    dw = (np.sin(T) * dtheta[:, None]) * (2 * math.pi / nphi)
    return sph(T, P), dw


D_vec = None


def D(h, ax, ay):
    return np.array([orc.ggx_D(x, ax, ay) for x in h.reshape(-1, 3)]).reshape(h.shape[:-1])


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.8, 0.8), (0.2, 0.6)])
def test_ggx_normalisation(ax, ay):
    """int D(h) cos(theta_h) dw = 1 (Trowbridge-Reitz, pbrt-v3 8.4)."""
    h, dw = hemisphere_grid()
    val = np.sum(D(h, ax, ay) * h[..., 2] * dw)
    assert abs(val - 1.0) < 3e-3, val


@pytest.mark.parametrize("ax,ay,theta_v,phi_v", [(0.3, 0.3, 0.6, 0.0), (0.5, 0.2, 1.0, 0.7), (0.9, 0.9, 1.3, 2.0)])
def test_visible_normal_identity(ax, ay, theta_v, phi_v):
    """int G1(v) max(0, v.h) D(h) dh = v.z pins D and Lambda together."""
    v = sph(np.array(theta_v), np.array(phi_v))
    h, dw = hemisphere_grid(400, 240)
    lam = orc.ggx_lambda(v, ax, ay)
    g1 = 1.0 / (1.0 + lam)
    val = g1 * np.sum(np.maximum(0.0, h @ v) * D(h, ax, ay) * dw)
    assert abs(val - v[2]) < 3e-3 * max(v[2], 0.2), (val, v[2])


def test_lambda_isotropic_closed_form_and_g1_bound():
    rng = np.random.default_rng(0)
    for _ in range(200):
        a = rng.uniform(0.05, 1.0)
        th = rng.uniform(0.0, 1.5)
        w = sph(np.array(th), np.array(rng.uniform(0, 2 * math.pi)))
        lam = orc.ggx_lambda(w, a, a)
        ref = (-1.0 + math.sqrt(1.0 + (a * math.tan(th)) ** 2)) / 2.0
        assert abs(lam - ref) <= 2e-5 * max(1.0, ref)
        assert 1.0 / (1.0 + lam) <= 1.0


def test_vndf_sampling_density():
    """Histogram of sampled cos(theta_h) against the VNDF D_v(h) = G1(v) max(0,v.h) D(h) / v.z."""
    ax, ay = 0.4, 0.4
    v = sph(np.array(0.8), np.array(0.3))
    rng = np.random.default_rng(1)
    n = 40000
    us = rng.random((n, 2))
    hs = np.array([orc.ggx_sample_vndf(v, ax, ay, a, b) for a, b in us])
    np.testing.assert_allclose(np.linalg.norm(hs, axis=1), 1.0, atol=1e-5)
    assert (hs @ v > -1e-6).all()                          # visible normals only
    bins = np.linspace(0.0, 1.0, 11)
    hist = np.histogram(hs[:, 2], bins=bins)[0] / n
    h, dw = hemisphere_grid(400, 240)
    g1 = 1.0 / (1.0 + orc.ggx_lambda(v, ax, ay))
    dens = g1 * np.maximum(0.0, h @ v) * D(h, ax, ay) / v[2] * dw
    expect = np.histogram(h[..., 2].ravel(), bins=bins, weights=dens.ravel())[0]
    assert abs(expect.sum() - 1.0) < 5e-3
    np.testing.assert_allclose(hist, expect, atol=0.006)


def material(diffuse, glossy, ax, ay):
    m = np.zeros((1, 36), np.float32)
    m[0, 4:7] = diffuse
    m[0, 8] = 1.0
    m[0, 12:15] = glossy
    m[0, 16] = 1.0
    m[0, 17] = ax
    m[0, 18] = ay
    return m


def setup(wi_world, n, mats):
    k = wi_world.shape[0]
    ray_d = np.zeros((k, 4), np.float32)
    ray_d[:, :3] = -wi_world
    nn = np.zeros((k, 4), np.float32)
    nn[:, :3] = n
    return orc.bsdf_setup(ray_d, nn, nn, np.zeros(k, np.int32), mats)


def test_lambert_closed_form():
    rng = np.random.default_rng(2)
    mats = material((0.6, 0.3, 0.9), (0.0, 0.0, 0.0), 0.5, 0.5)
    k = 500
    wi = sph(rng.uniform(0, 1.4, k), rng.uniform(0, 6.28, k))
    wo = sph(rng.uniform(0, 1.4, k), rng.uniform(0, 6.28, k))
    rec, mid = setup(wi.astype(np.float32), np.tile([0, 0, 1], (k, 1)), mats)
    assert (rec[:, 15] == 1.0).all() and (rec[:, 16] == 0.0).all()   # diffProb 1, glossyProb 0
    f, pdf, cs = orc.bsdf_eval(rec, mid, mats, wo.astype(np.float32))
    np.testing.assert_allclose(f, np.tile(np.array([0.6, 0.3, 0.9]) / math.pi, (k, 1)), rtol=1e-6)
    np.testing.assert_allclose(pdf, wo[:, 2] / math.pi, rtol=1e-5)
    np.testing.assert_allclose(cs, wo[:, 2], rtol=1e-5)


def test_glossy_reciprocity():
    """f(wi -> wo) == f(wo -> wi): G, D(wh) and Schlick(|wh.wo| == |wh.wi|) are symmetric."""
    rng = np.random.default_rng(3)
    mats = material((0.5, 0.5, 0.5), (0.04, 0.04, 0.04), 0.3, 0.7)
    k = 300
    n = np.tile([0.0, 0.0, 1.0], (k, 1))
    wi = sph(rng.uniform(0.1, 1.3, k), rng.uniform(0, 6.28, k)).astype(np.float32)
    wo = sph(rng.uniform(0.1, 1.3, k), rng.uniform(0, 6.28, k)).astype(np.float32)
    r1, m1 = setup(wi, n, mats)
    r2, m2 = setup(wo, n, mats)
    f1, _, _ = orc.bsdf_eval(r1, m1, mats, wo)
    f2, _, _ = orc.bsdf_eval(r2, m2, mats, wi)
    np.testing.assert_allclose(f1, f2, rtol=2e-5)


def test_sample_matches_eval_of_sampled_direction():
    """f returned by SampleBsdf equals EvaluateBsdf at the sampled direction."""
    rng = np.random.default_rng(4)
    mats = material((0.5, 0.4, 0.3), (0.04, 0.04, 0.04), 0.4, 0.4)
    k = 400
    wi = sph(rng.uniform(0.1, 1.2, k), rng.uniform(0, 6.28, k)).astype(np.float32)
    rec, mid = setup(wi, np.tile([0.0, 0.0, 1.0], (k, 1)), mats)
    u = rng.random((k, 3)).astype(np.float32)
    f, d, pdf, cs = orc.bsdf_sample(rec, mid, mats, u)
    ok = (f != 0).any(1) & (d[:, 2] > 1e-3)
    fe, pe, ce = orc.bsdf_eval(rec[ok], mid[ok], mats, d[ok])
    np.testing.assert_allclose(fe, f[ok], rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(np.linalg.norm(d[ok], axis=1), 1.0, atol=1e-5)


def _wang_hash_py(seed):
    """Independent restatement with explicit 32-bit wrap and arithmetic shifts."""
    def i32(x):
        x &= 0xFFFFFFFF
        return x - (1 << 32) if x & 0x80000000 else x
    s = i32((seed ^ 61) ^ (i32(seed) >> 16))
    s = i32(s * 9)
    s = i32(s ^ (s >> 4))
    s = i32(s * 0x27D4EB2D)
    s = i32(s ^ (s >> 15))
    return s & 0xFFFFFFFF


def test_wang_hash_known_answers():
    # hand-derived: seed 0 -> 61 -> 549 -> 549^34 = 519 -> 519*0x27d4eb2d mod 2^32
    assert _wang_hash_py(0) == orc.wang_hash(0)
    s = (519 * 0x27D4EB2D) & 0xFFFFFFFF
    s_signed = s - (1 << 32)
    assert orc.wang_hash(0) == (s_signed ^ (s_signed >> 15)) & 0xFFFFFFFF
    for seed in [1, 2, 12345, -1, -987654321, 2 ** 31 - 1, -2 ** 31]:
        assert orc.wang_hash(seed) == _wang_hash_py(seed), seed


def test_get_random_range_and_uniformity():
    st = np.array([12345, 67890], np.int32)
    xs = []
    for _ in range(20000):
        r, st = orc.get_random(st)
        xs.append(r)
    xs = np.array(xs)
    assert (xs >= 0).all() and (xs < 1).all()
    hist = np.histogram(xs, bins=20, range=(0, 1))[0]
    chi2 = np.sum((hist - 1000.0) ** 2 / 1000.0)
    assert chi2 < 60, chi2       # 19 dof; p ~ 1e-6 at 60
