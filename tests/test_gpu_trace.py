"""Stage A on the GPU against the oracle: closest hit, any hit, traversal
statistics, hit shading and shadow rays, in both memory layouts.

Bar: primitive ids and occlusion flags bit-exact, t bit-exact (same
evaluation order, no FMA, IEEE division), per-ray interior/leaf/pop counts
equal (same visiting order).
"""
import numpy as np
import pytest
import torch

from mambatracer_amd import api
from oracle import orc_py as orc
from tests.conftest import c0_rays

pytestmark = pytest.mark.gpu

MODES = [api._lib.MT_TRACE_PACKED, api._lib.MT_TRACE_REFERENCE, api._lib.MT_TRACE_PACKED_NO_STACK,
         api._lib.MT_TRACE_PACKED_STACK, api._lib.MT_TRACE_PACKED_REFILL, api._lib.MT_TRACE_PACKED_SORTED,
         api._lib.MT_TRACE_PACKED_POOL, api._lib.MT_TRACE_PACKED_BINNED]


def _scene(ctx, arrs, normals=None):
    points, faces, nodes, bounds, depth = arrs
    return api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)


@pytest.mark.parametrize("mode", MODES)
def test_c0_closest_hit_bit_exact(ctx, c0, mode):
    print(f"[case] c0 closest mode={mode}", flush=True)
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays()
    sc = _scene(ctx, c0)
    prim, t, st = api.intersect_primitives(ctx, sc, ctx.to_device(o), ctx.to_device(d), mode=mode, stats=True)
    torch.cuda.synchronize()
    oprim, ot, otot, ors = orc.trace_closest(points, faces, nodes, bounds, o, d, ray_stats=True)
    np.testing.assert_array_equal(prim.cpu().numpy(), oprim)
    np.testing.assert_array_equal(t.cpu().numpy().view(np.int32), ot.view(np.int32))
    np.testing.assert_array_equal(st.cpu().numpy(), ors)


@pytest.mark.parametrize("mode", MODES)
def test_c0_any_hit(ctx, c0, mode):
    print(f"[case] c0 any mode={mode}", flush=True)
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(256, 256)
    # shorten tMax so that some rays are unoccluded
    rng = np.random.default_rng(5)
    d = d.copy()
    d[:, 3] = rng.uniform(0.0, 2.0, size=d.shape[0]).astype(np.float32)
    sc = _scene(ctx, c0)
    occ, st = api.intersect_occlusion(ctx, sc, ctx.to_device(o), ctx.to_device(d), mode=mode, stats=True)
    torch.cuda.synchronize()
    oocc, _, ors = orc.trace_any(points, faces, nodes, bounds, o, d, ray_stats=True)
    np.testing.assert_array_equal(occ.cpu().numpy(), oocc)
    np.testing.assert_array_equal(st.cpu().numpy(), ors)
    assert 0 < oocc.sum() < len(oocc)


def test_device_count_guard(ctx, c0):
    """Slots >= *count are untouched (IntersectPrimitives' `id < *count`, NormalBVH.cl:120)."""
    print("[case] count guard", flush=True)
    o, d = c0_rays(64, 64)
    sc = _scene(ctx, c0)
    n = o.shape[0]
    count = ctx.scalar_i32(1000)
    prim = torch.full((n,), -7, dtype=torch.int32, device=ctx.torch_device)
    t = torch.full((n,), -7.0, dtype=torch.float32, device=ctx.torch_device)
    api.call("mt_trace_closest", ctx.handle, sc.handle, ctx.to_device(o).data_ptr(), ctx.to_device(d).data_ptr(),
             count.data_ptr(), n, prim.data_ptr(), t.data_ptr(), None, 0)
    p = prim.cpu().numpy()
    assert (p[1000:] == -7).all() and (p[:1000] != -7).all()


def test_random_rays_small_scene_and_deep_tree(ctx):
    """Stress soup with coincident centroids (deep tree) and random rays, seed 3 layout."""
    print("[case] stress small", flush=True)
    tp, tf = api.gen_soup(3, 30000, -0.05, 0.05, 1e-4, 1e-3, first_vertex=0)
    lp, lf = api.gen_soup(4, 3000, -100.0, 100.0, 1.0, 10.0, first_vertex=tp.shape[0])
    points = np.concatenate([tp, lp])
    faces = np.concatenate([tf, lf])
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    lo, hi = bounds[0, :3], bounds[0, 4:7]
    o, d = api.gen_random_rays(7, 0, 100000, lo, hi)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth)
    for mode in MODES:
        prim, t, st = api.intersect_primitives(ctx, sc, ctx.to_device(o), ctx.to_device(d), mode=mode, stats=True)
        oprim, ot, _, ors = orc.trace_closest(points, faces, nodes, bounds, o, d, ray_stats=True)
        np.testing.assert_array_equal(prim.cpu().numpy(), oprim)
        np.testing.assert_array_equal(t.cpu().numpy().view(np.int32), ot.view(np.int32))
        np.testing.assert_array_equal(st.cpu().numpy(), ors)


def test_shade_hits(ctx):
    print("[case] shading", flush=True)
    points, normals, faces = api.gen_spheres(2, 4)
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    o, d = api.gen_camera_rays([0, 1, -12], [0, 0, 0], [0, 1, 0], 60.0, 128, 128)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    prim, t, _ = api.intersect_primitives(ctx, sc, do, dd)
    p, n, ng, uv, mat = api.shade_hits(ctx, sc, do, dd, prim, t)
    torch.cuda.synchronize()
    op, on, ong, ouv, omat = orc.shade_hits(points, normals, None, faces, o, d, prim.cpu().numpy(), t.cpu().numpy())
    np.testing.assert_array_equal(mat.cpu().numpy(), omat)
    for a, b in [(p, op), (n, on), (ng, ong), (uv, ouv)]:
        np.testing.assert_array_equal(a.cpu().numpy().view(np.int32), b.view(np.int32))
    assert (omat >= 0).sum() > 1000


def test_shadow_rays_and_compaction_pipeline(ctx, c0):
    """Config-2 pipeline shape on the c0 scene: closest hit -> shadow rays ->
    any hit -> stable compaction of unoccluded hits, all against the oracle."""
    print("[case] shadow pipeline", flush=True)
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(256, 256)
    sc = _scene(ctx, c0)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    prim, t, _ = api.intersect_primitives(ctx, sc, do, dd)
    light = np.array([0.3, 0.9, 0.2], np.float32)
    so, sd, valid = api.shadow_rays(ctx, do, dd, prim, t, light)
    occ, _ = api.intersect_occlusion(ctx, sc, so, sd)
    lit = (valid == 1) & (occ == 0)
    idx, cnt, _ = api.compact(ctx, lit.to(torch.int32))
    torch.cuda.synchronize()
    oprim, ot, _, _ = orc.trace_closest(points, faces, nodes, bounds, o, d)
    oso, osd, ovalid = orc.shadow_rays(o, d, oprim, ot, light)
    np.testing.assert_array_equal(so.cpu().numpy().view(np.int32), oso.view(np.int32))
    np.testing.assert_array_equal(sd.cpu().numpy().view(np.int32), osd.view(np.int32))
    oocc, _, _ = orc.trace_any(points, faces, nodes, bounds, oso, osd)
    olit = (ovalid == 1) & (oocc == 0)
    oidx, ocnt = orc.compact_stable(olit.astype(np.int32))
    assert int(cnt.item()) == ocnt
    np.testing.assert_array_equal(idx.cpu().numpy()[:ocnt], oidx)
    ctx.check_scans()


@pytest.mark.parametrize("any_hit", [False, True])
def test_autotune_picks_a_packed_mode_with_identical_results(ctx, c0, any_hit):
    """mt_trace_autotune times the short-stack and refill modes on the batch;
    whichever it picks, its outputs equal the oracle's (every mode is exact)."""
    print(f"[case] autotune any={any_hit}", flush=True)
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(256, 256)
    sc = _scene(ctx, c0)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    mode, ms = api.autotune_trace(ctx, sc, do, dd, any_hit=any_hit, reps=2)
    assert mode in (api._lib.MT_TRACE_PACKED_STACK, api._lib.MT_TRACE_PACKED_REFILL) and ms > 0
    if any_hit:
        occ, _ = api.intersect_occlusion(ctx, sc, do, dd, mode=mode)
        oocc, _, _ = orc.trace_any(points, faces, nodes, bounds, o, d)
        np.testing.assert_array_equal(occ.cpu().numpy(), oocc)
    else:
        prim, t, _ = api.intersect_primitives(ctx, sc, do, dd, mode=mode)
        oprim, ot, _, _ = orc.trace_closest(points, faces, nodes, bounds, o, d)
        np.testing.assert_array_equal(prim.cpu().numpy(), oprim)
        np.testing.assert_array_equal(t.cpu().numpy().view(np.int32), ot.view(np.int32))
    with pytest.raises(RuntimeError, match="reps"):
        api.autotune_trace(ctx, sc, do, dd, any_hit=any_hit, reps=0)


@pytest.mark.parametrize("mode", [api._lib.MT_TRACE_PACKED, api._lib.MT_TRACE_REFERENCE,
                                  api._lib.MT_TRACE_PACKED_NO_STACK])
def test_intersect_records_against_oracle(ctx, mode):
    """IntersectPrimitives over the reference's 80-byte Ray and 96-byte
    Intersection records: hits bit-exact against the oracle's closest hit and
    hit shading (normals on every face, UVs on every other face), tMax
    written back on hits only, misses and slots past *count writing only
    hit = 0 and mat = -1."""
    print(f"[case] intersection records mode={mode}", flush=True)
    points, normals, faces = api.gen_spheres(2, 4)
    faces = faces.copy()
    faces[0::2, 3:6] = faces[0::2, 0:3]
    uvs = np.random.default_rng(5).random((points.shape[0], 2), dtype=np.float32)
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    o, d = api.gen_camera_rays([0, 1, -12], [0, 0, 0], [0, 1, 0], 60.0, 128, 128)
    n = o.shape[0]
    count = n - 100
    rec = api.ray_records(o, d)
    rec[:, 7] = np.float32(0.25)                       # d.w travels into the intersection
    rec.view(np.int32)[:, 16:18] = np.arange(2 * n, dtype=np.int32).reshape(n, 2)  # extra untouched
    sentinel = np.full((n, api.INTERSECTION_RECORD_WORDS), 0x7F7F7F7F, np.int32)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals, uvs=uvs)
    drays, disect = ctx.to_device(rec), ctx.to_device(sentinel)
    api.intersect_records(ctx, sc, drays, disect, count=ctx.scalar_i32(count), mode=mode)
    torch.cuda.synchronize()
    got_r, got_i = drays.cpu().numpy(), disect.cpu().numpy()
    oprim, ot, _, _ = orc.trace_closest(points, faces, nodes, bounds, o[:count], d[:count])
    op, on, ong, ouv, omat = orc.shade_hits(points, normals, uvs, faces, o[:count], d[:count], oprim, ot)
    hit = np.zeros(n, bool)
    hit[:count] = oprim >= 0
    assert hit.sum() > 1000 and (~hit[:count]).sum() > 100 and (faces[oprim[oprim >= 0], 3] >= 0).any()
    # rays: only tMax of hit slots changes, to the closest hit's distance
    want_r = rec.copy()
    want_r[:count, 19] = np.where(oprim >= 0, ot, rec[:count, 19])
    np.testing.assert_array_equal(got_r.view(np.int32), want_r.view(np.int32))
    gi = got_i
    gf = got_i.view(np.float32)
    h = np.nonzero(hit)[0]
    np.testing.assert_array_equal(gi[h, 20], 1)
    np.testing.assert_array_equal(gi[h, 19], oprim[h])
    np.testing.assert_array_equal(gi[h, 18], omat[h])
    for lo, want in [(0, op), (4, on), (8, ong)]:
        np.testing.assert_array_equal(gf[h, lo:lo + 4].view(np.int32), want[h].view(np.int32))
    np.testing.assert_array_equal(gf[h, 12:16].view(np.int32), rec[h, 4:8].view(np.int32))
    np.testing.assert_array_equal(gf[h, 16:18].view(np.int32), ouv[h].view(np.int32))
    np.testing.assert_array_equal(gi[h, 21:24], 0x7F7F7F7F)
    m = np.nonzero(~hit)[0]
    np.testing.assert_array_equal(gi[m, 20], 0)
    np.testing.assert_array_equal(gi[m, 18], -1)
    keep = np.r_[0:18, 19, 21:24]
    np.testing.assert_array_equal(gi[m][:, keep], 0x7F7F7F7F)


@pytest.mark.parametrize("mode", [api._lib.MT_TRACE_PACKED, api._lib.MT_TRACE_REFERENCE,
                                  api._lib.MT_TRACE_PACKED_NO_STACK])
def test_occlusion_records_against_oracle(ctx, c0, mode):
    """intersectOcclusion over 80-byte Ray records: occlusion bit-exact
    against the oracle's any hit, rays unchanged, slots past *count left as
    they were."""
    print(f"[case] occlusion records mode={mode}", flush=True)
    points, faces, nodes, bounds, _ = c0
    o, d = c0_rays(256, 256)
    sc = _scene(ctx, c0)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    prim, t, _ = api.intersect_primitives(ctx, sc, do, dd)
    so, sd, _ = api.shadow_rays(ctx, do, dd, prim, t, np.array([0.3, 0.9, 0.2], np.float32))
    so, sd = so.cpu().numpy(), sd.cpu().numpy()
    n = so.shape[0]
    count = n - 77
    rec = api.ray_records(so, sd)
    drays = ctx.to_device(rec)
    hits = ctx.to_device(np.full(n, -5, np.int32))
    api.intersect_occlusion_records(ctx, sc, drays, hits, count=ctx.scalar_i32(count), mode=mode)
    torch.cuda.synchronize()
    oocc, _, _ = orc.trace_any(points, faces, nodes, bounds, so[:count], sd[:count])
    got = hits.cpu().numpy()
    np.testing.assert_array_equal(got[:count], oocc)
    np.testing.assert_array_equal(got[count:], -5)
    np.testing.assert_array_equal(drays.cpu().numpy().view(np.int32), rec.view(np.int32))
    assert 0 < oocc.sum() < count


def test_records_edge_cases(ctx):
    """Record entry points on a zero count (every slot a miss, nothing else
    written) and on a one-triangle scene (root is a leaf)."""
    print("[case] record edge cases", flush=True)
    tri = np.array([[0, 0, 1, 0], [1, 0, 1, 0], [0, 1, 1, 0]], np.float32)
    faces = np.array([[0, 1, 2, -1, -1, -1, -1, -1, -1, 7]], np.int32)
    nodes, bounds, depth = api.CNormalBVH.build(tri, faces)
    sc = api.Scene(ctx, tri, faces, nodes, bounds, depth)
    o = np.array([[0.2, 0.2, 0, 0.001], [0.9, 0.9, 0, 0.001], [0.2, 0.2, 2, 0.001]], np.float32)
    d = np.array([[0, 0, 1, np.inf], [0, 0, 1, np.inf], [0, 0, 1, np.inf]], np.float32)
    rec = api.ray_records(o, d)
    sentinel = np.full((3, api.INTERSECTION_RECORD_WORDS), 0x55, np.int32)
    for count in (0, 3):
        drays, dis = ctx.to_device(rec), ctx.to_device(sentinel)
        api.intersect_records(ctx, sc, drays, dis, count=ctx.scalar_i32(count))
        hits = ctx.to_device(np.full(3, -5, np.int32))
        api.intersect_occlusion_records(ctx, sc, drays, hits, count=ctx.scalar_i32(count))
        torch.cuda.synchronize()
        gi, gr = dis.cpu().numpy(), drays.cpu().numpy()
        if count == 0:
            np.testing.assert_array_equal(gi[:, 20], 0)
            np.testing.assert_array_equal(gi[:, 18], -1)
            np.testing.assert_array_equal(gr.view(np.int32), rec.view(np.int32))
            np.testing.assert_array_equal(hits.cpu().numpy(), -5)
        else:
            np.testing.assert_array_equal(gi[:, 20], [1, 0, 0])   # inside, outside the triangle, behind the origin
            assert gi[0, 19] == 0 and gi[0, 18] == 7 and gr.view(np.float32)[0, 19] == np.float32(1.0)
            np.testing.assert_array_equal(gi[1:, 18], -1)
            np.testing.assert_array_equal(gr[1:].view(np.int32), rec[1:].view(np.int32))
            # the first ray's tMax is now 1: the occlusion pass sees the hit at t = 1 as outside (tMin, tMax)
            np.testing.assert_array_equal(hits.cpu().numpy(), [0, 0, 0])


@pytest.mark.parametrize("mode", [api._lib.MT_TRACE_PACKED_POOL, api._lib.MT_TRACE_PACKED_BINNED])
def test_sorted_and_pooled_modes_at_edge_counts(ctx, c0, mode):
    """Live counts of 0 and 1 and around the binned sort's 65,536-ray tile
    (one tile minus one, exactly one, one more, two partial tiles): every
    live ray's result equals one ray per lane."""
    print(f"[case] edge counts mode={mode}", flush=True)
    sc = _scene(ctx, c0)
    o, d = api.gen_random_rays(91, 0, 70_000, [-1.0] * 3, [1.0] * 3)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    for n in (0, 1, 65_535, 65_536, 65_537, 70_000):
        count = ctx.scalar_i32(n)
        for fn in (api.intersect_primitives, api.intersect_occlusion):
            a = fn(ctx, sc, do, dd, count=count, mode=api._lib.MT_TRACE_PACKED_STACK)
            b = fn(ctx, sc, do, dd, count=count, mode=mode)
            torch.cuda.synchronize()
            for x, y in zip(a, b):
                if x is None:
                    continue
                assert np.array_equal(x.cpu().numpy()[:n].view(np.int32), y.cpu().numpy()[:n].view(np.int32)), n


def test_binned_mode_on_a_flat_scene_and_non_finite_origins(ctx):
    """The binned sort's keys on a root box with a zero extent (every
    triangle in the plane z = 0, so that axis has no cells) and on origins at
    +-inf or NaN (clamped to the first or last cell): each live ray's closest
    and any hit equal one ray per lane."""
    print("[case] binned mode, flat scene", flush=True)
    points, faces = api.gen_soup(17, 5000, -1.0, 1.0, 0.01, 0.2, first_vertex=0)
    points[:, 2] = 0.0
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth)
    n = 70_000
    o, d = api.gen_random_rays(95, 0, n, [-1.2, -1.2, -1.0], [1.2, 1.2, 1.0])
    o[:50, 0] = np.inf
    o[50:100, 1] = -np.inf
    o[100:150, 2] = np.nan
    o[150:200, 2] = 0.0  # origins in the scene's plane
    do, dd = ctx.to_device(o), ctx.to_device(d)
    count = ctx.scalar_i32(n - 3)
    for fn in (api.intersect_primitives, api.intersect_occlusion):
        a = fn(ctx, sc, do, dd, count=count, mode=api._lib.MT_TRACE_PACKED_STACK)
        b = fn(ctx, sc, do, dd, count=count, mode=api._lib.MT_TRACE_PACKED_BINNED)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            if x is None:
                continue
            assert torch.equal(x[:n - 3].view(torch.int32), y[:n - 3].view(torch.int32))


def test_binned_mode_with_batch_sized_tiles(ctx, c0):
    """A batch large enough that the binned sort's tiles grow past their
    65,536-ray minimum (20,000,003 rays: 254 tiles of 78,848, the last one
    partial) with a live count that ends inside the last tile: every live
    ray's closest and any hit equal one ray per lane."""
    print("[case] binned mode, 20M rays", flush=True)
    sc = _scene(ctx, c0)
    n = 20_000_003
    o, d = api.gen_random_rays(93, 0, n, [-1.0] * 3, [1.0] * 3)
    do, dd = ctx.to_device(o), ctx.to_device(d)
    del o, d
    count = ctx.scalar_i32(n - 5)
    for fn in (api.intersect_primitives, api.intersect_occlusion):
        a = fn(ctx, sc, do, dd, count=count, mode=api._lib.MT_TRACE_PACKED_STACK)
        b = fn(ctx, sc, do, dd, count=count, mode=api._lib.MT_TRACE_PACKED_BINNED)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            if x is None:
                continue
            assert torch.equal(x[:n - 5].view(torch.int32), y[:n - 5].view(torch.int32))


@pytest.mark.parametrize("mode", [api._lib.MT_TRACE_PACKED_SORTED, api._lib.MT_TRACE_PACKED_POOL,
                                  api._lib.MT_TRACE_PACKED_BINNED])
@pytest.mark.parametrize("any_hit", [False, True])
def test_lane_mappings_same_results(ctx, c0, any_hit, mode):
    """MT_TRACE_PACKED_SORTED (each work-group's rays taken in direction-bin
    order), MT_TRACE_PACKED_POOL (each wave refilling its finished lanes
    from its own run of rays) and MT_TRACE_PACKED_BINNED (the batch sorted by
    origin cell first) return exactly the per-ray
    results of one ray per lane, on camera rays and on random rays with a
    live length that ends mid-group and mid-run."""
    print(f"[case] lane mapping mode={mode} any={any_hit}", flush=True)
    points, faces, nodes, bounds, _ = c0
    sc = _scene(ctx, c0)
    o1, d1 = c0_rays(128, 128)
    o2, d2 = api.gen_random_rays(77, 0, 50_000, [-1.0] * 3, [1.0] * 3)
    for o, d in ((o1, d1), (o2, d2)):
        do, dd = ctx.to_device(o), ctx.to_device(d)
        count = ctx.scalar_i32(len(o) - 37)
        fn = api.intersect_occlusion if any_hit else api.intersect_primitives
        a = fn(ctx, sc, do, dd, count=count, mode=api._lib.MT_TRACE_PACKED_STACK)
        b = fn(ctx, sc, do, dd, count=count, mode=mode)
        torch.cuda.synchronize()
        n = len(o) - 37
        for x, y in zip(a, b):
            if x is None:
                continue
            xa, ya = x.cpu().numpy()[:n], y.cpu().numpy()[:n]
            assert np.array_equal(xa.view(np.int32), ya.view(np.int32))
