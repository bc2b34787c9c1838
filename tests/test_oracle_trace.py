"""Pins the oracle's BVH builder and traversal without the reference:
tree invariants, product builder == oracle builder, brute force over all
triangles, any-hit consistency, analytic room-box distances, and
barycentric reconstruction of shaded hit points."""
import numpy as np
import pytest

from mambatracer_amd import api
from oracle import orc_py as orc
from tests.conftest import c0_rays


def check_tree(nodes, bounds, points, faces):
    n = faces.shape[0]
    assert nodes.shape == (2 * n - 1, 7)
    assert (nodes[:, 0] == np.arange(2 * n - 1)).all()              # bound == own index
    leaf = nodes[:, 5] == 1
    assert leaf.sum() == n
    assert sorted(nodes[leaf, 6].tolist()) == list(range(n))          # leaves: a permutation
    inner = np.flatnonzero(~leaf)
    L, R = nodes[inner, 2], nodes[inner, 3]
    assert (nodes[L, 4] == inner).all() and (nodes[R, 4] == inner).all()
    assert (nodes[L, 1] == R).all() and (nodes[R, 1] == L).all()     # sibling links
    for ch in (L, R):                                                 # boxes nest
        assert (bounds[ch, :3] >= bounds[inner, :3]).all()
        assert (bounds[ch, 4:7] <= bounds[inner, 4:7]).all()
    li = np.flatnonzero(leaf)
    tri = points[faces[nodes[li, 6], :3], :3]                         # leaf box = triangle box
    np.testing.assert_array_equal(bounds[li, :3], tri.min(1))
    np.testing.assert_array_equal(bounds[li, 4:7], tri.max(1))


def test_builder_invariants_and_host_product_equal(c0):
    points, faces, nodes, bounds, depth = c0
    check_tree(nodes, bounds, points, faces)
    on, ob, od = orc.bvh_build(points, faces)
    np.testing.assert_array_equal(on, nodes)
    np.testing.assert_array_equal(ob.view(np.int32), bounds.view(np.int32))
    assert od == depth <= 64


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100])
def test_builder_small(n):
    p, f = api.gen_soup(9, n, -1, 1, 0.1, 0.5)
    nodes, bounds, depth = api.CNormalBVH.build(p, f)
    check_tree(nodes, bounds, p, f)
    on, ob, _ = orc.bvh_build(p, f)
    np.testing.assert_array_equal(on, nodes)


def test_builder_coincident_centroids():
    """All centroids equal: every partition is degenerate -> count-median split
    (CNormalBVH.java:135-136)."""
    p1, f1 = api.gen_soup(5, 1, 0.0, 0.0, 0.1, 0.2)
    f = np.repeat(f1, 64, axis=0)          # 64 copies of one triangle
    p = p1
    nodes, bounds, depth = api.CNormalBVH.build(p, f)
    check_tree(nodes, bounds, p, f)
    assert depth == 6
    on, _, _ = orc.bvh_build(p, f)
    np.testing.assert_array_equal(on, nodes)


def test_closest_hit_against_brute_force(c0):
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(128, 128)
    prim, t, _, _ = orc.trace_closest(points, faces, nodes, bounds, o, d)
    sel = np.random.default_rng(1).choice(len(o), 3000, replace=False)
    bp, bt = orc.trace_brute(points, faces, o[sel], d[sel])
    np.testing.assert_array_equal(bt.view(np.int32), t[sel].view(np.int32))
    diff = np.flatnonzero(bp != prim[sel])
    # the only admissible id differences are exact t-ties between two triangles
    for k in diff:
        _, t_other = orc.trace_brute(points, faces[[prim[sel][k]]], o[sel][k:k + 1], d[sel][k:k + 1])
        assert t_other[0] == bt[k]
    print(f"exact t-ties with a different id (shared room edges): {len(diff)} of {len(sel)}")
    assert len(diff) <= len(sel) // 200


def test_random_rays_brute_force():
    p, f = api.gen_soup(11, 3000, -1, 1, 0.01, 0.3)
    nodes, bounds, _ = api.CNormalBVH.build(p, f)
    o, d = api.gen_random_rays(12, 0, 2000, [-1.2] * 3, [1.2] * 3)
    prim, t, _, _ = orc.trace_closest(p, f, nodes, bounds, o, d)
    bp, bt = orc.trace_brute(p, f, o, d)
    np.testing.assert_array_equal(bt.view(np.int32), t.view(np.int32))
    # every id difference must be an exact t-tie: the traversal's triangle
    # reaches the brute-force distance too (listed)
    diff = np.flatnonzero(bp != prim)
    for k in diff:
        _, t_other = orc.trace_brute(p, f[[prim[k]]], o[k:k + 1], d[k:k + 1])
        print(f"ray {k}: brute {bp[k]} vs traversal {prim[k]}, t {bt[k]!r} both")
        assert t_other[0] == bt[k] and prim[k] >= 0
    assert len(diff) <= 2


def test_any_hit_equals_closest_within_tmax(c0):
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(96, 96)
    d = d.copy()
    d[:, 3] = np.random.default_rng(3).uniform(0, 2.5, size=len(d)).astype(np.float32)
    prim, _, _, _ = orc.trace_closest(points, faces, nodes, bounds, o, d)
    occ, _, _ = orc.trace_any(points, faces, nodes, bounds, o, d)
    np.testing.assert_array_equal(occ, (prim >= 0).astype(np.int32))


def test_room_box_analytic_distance():
    """Empty room [-1,1]^3: t is the exit distance of the ray from the box."""
    rp, rf = api.gen_room(0)
    nodes, bounds, _ = api.CNormalBVH.build(rp, rf)
    o, d = api.gen_random_rays(4, 0, 5000, [-0.9] * 3, [0.9] * 3)
    prim, t, _, _ = orc.trace_closest(rp, rf, nodes, bounds, o, d)
    O, D = o[:, :3].astype(np.float64), d[:, :3].astype(np.float64)
    with np.errstate(divide="ignore"):
        texit = np.min(np.where(D != 0, (np.sign(D) - O) / D, np.inf), axis=1)
    assert (prim >= 0).mean() > 0.995
    ok = prim >= 0
    np.testing.assert_allclose(t[ok], texit[ok], rtol=2e-6)
    # the hit wall is the one whose plane the exit point lies on
    axis = np.argmin(np.where(D != 0, (np.sign(D) - O) / D, np.inf), axis=1)
    for k in np.flatnonzero(ok)[:500]:
        a, s = axis[k], np.sign(D[k, axis[k]])
        base = {(0, -1): 8, (0, 1): 10, (1, -1): 4, (1, 1): 6, (2, -1): 0, (2, 1): 2}[(a, int(s))]
        assert prim[k] in (base, base + 1)


def test_shading_reconstructs_hit_point(c0):
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(64, 64)
    prim, t, _, _ = orc.trace_closest(points, faces, nodes, bounds, o, d)
    p, n, ng, uv, mat = orc.shade_hits(points, None, None, faces, o, d, prim, t)
    ok = prim >= 0
    P = points[faces[prim[ok], :3], :3].astype(np.float64)
    u, v = uv[ok, 0].astype(np.float64), uv[ok, 1].astype(np.float64)
    rec = (1 - u - v)[:, None] * P[:, 0] + u[:, None] * P[:, 1] + v[:, None] * P[:, 2]
    np.testing.assert_allclose(rec, p[ok, :3], atol=2e-5)
    np.testing.assert_allclose(np.linalg.norm(n[ok, :3], axis=1), 1.0, atol=1e-6)
    assert (np.sum(n[ok, :3] * d[ok, :3], 1) <= 1e-5).all()          # faces the viewer
    assert (mat[ok] == 0).all() and (mat[~ok] == -1).all()


def test_shadow_rays_point_at_light(c0):
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(32, 32)
    prim, t, _, _ = orc.trace_closest(points, faces, nodes, bounds, o, d)
    light = np.array([0.2, 0.5, 0.1], np.float32)
    so, sd, valid = orc.shadow_rays(o, d, prim, t, light)
    ok = valid == 1
    dist = np.linalg.norm(light - so[ok, :3], axis=1)
    np.testing.assert_allclose(sd[ok, 3], dist - 0.002, rtol=1e-5)
    np.testing.assert_allclose(so[ok, :3] + sd[ok, :3] * dist[:, None], np.broadcast_to(light, (ok.sum(), 3)),
                               atol=1e-5)
    assert (so[:, 3] == np.float32(0.0001)).all()
