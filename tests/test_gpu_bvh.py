"""Device BVH builder (mt_bvh_build_device) against the host builder and the
oracle's builder (CNormalBVH.build restated, oracle/orc_trace.c).

Bar: node records bit-exact, bounds equal as floats (a zero bound may carry
the other sign), same depth; the cases cover the fallback split (all
centroids equal: the range is halved by index), leaves of one primitive,
one- to three-triangle scenes and scenes deep enough to exercise many
levels.  A scene built on the device traces exactly like the host-built one."""
import numpy as np
import pytest
import torch

from mambatracer_amd import api
from mambatracer_amd import workloads as W
from oracle import orc_py as orc
from tests.conftest import c0_rays

pytestmark = pytest.mark.gpu


def _cases():
    out = []
    one = api.gen_soup(11, 1, -1, 1, 0.1, 0.2)
    out.append(("one", *one))
    out.append(("two", *api.gen_soup(12, 2, -1, 1, 0.1, 0.2)))
    out.append(("three", *api.gen_soup(13, 3, -1, 1, 0.1, 0.2)))
    p, f = api.gen_soup(14, 1, -1, 1, 0.1, 0.2)
    out.append(("coincident64", p, np.repeat(f, 64, axis=0)))
    tp, tf = api.gen_soup(15, 5000, -0.01, 0.01, 1e-4, 1e-3)
    dp, df = api.gen_soup(16, 1, -1, 1, 0.1, 0.2, first_vertex=tp.shape[0])
    out.append(("clustered+dups", np.concatenate([tp, dp]), np.concatenate([tf, np.repeat(df, 300, axis=0)])))
    out.append(("soup100k", *api.gen_soup(17, 100_000, -5, 5, 0.01, 0.5)))
    pts, _, fc = api.gen_spheres(2, 8)
    out.append(("spheres", pts, fc))
    return out


@pytest.mark.parametrize("case", range(7))
def test_device_builder_matches_host_and_oracle(ctx, case):
    name, points, faces = _cases()[case]
    print(f"[case] device bvh {name} n={faces.shape[0]}", flush=True)
    hn, hb, hd = api.CNormalBVH.build(points, faces)
    dn, db, dd = api.CNormalBVH.build_device(ctx, ctx.to_device(points), ctx.to_device(faces))
    torch.cuda.synchronize()
    dn, db = dn.cpu().numpy(), db.cpu().numpy()
    np.testing.assert_array_equal(dn, hn)
    np.testing.assert_array_equal(db, hb)   # == treats -0.0 and 0.0 as equal
    assert dd == hd
    on, ob, od = orc.bvh_build(points, faces)
    np.testing.assert_array_equal(dn, on)
    np.testing.assert_array_equal(db, ob)
    assert dd == od


def test_device_built_scene_traces_like_host_built(ctx, c0):
    points, faces, nodes, bounds, depth = c0
    dp, df = ctx.to_device(points), ctx.to_device(faces)
    dn, db, dd = api.CNormalBVH.build_device(ctx, dp, df)
    assert dd == depth
    sc_h = api.Scene(ctx, points, faces, nodes, bounds, depth)
    sc_d = api.Scene(ctx, points, faces, dn.cpu().numpy(), db.cpu().numpy(), dd)
    o, d = c0_rays(128, 128)
    do, dd_ = ctx.to_device(o), ctx.to_device(d)
    p1, t1, _ = api.intersect_primitives(ctx, sc_h, do, dd_)
    p2, t2, _ = api.intersect_primitives(ctx, sc_d, do, dd_)
    torch.cuda.synchronize()
    assert torch.equal(p1, p2) and torch.equal(t1.view(torch.int32), t2.view(torch.int32))


def test_device_builder_large(ctx):
    """1M-triangle c1 scene: same tree as the host builder."""
    print("[case] device bvh c1", flush=True)
    points, normals, faces = api.gen_spheres(2, 256)
    hn, hb, hd = api.CNormalBVH.build(points, faces)
    dn, db, dd = api.CNormalBVH.build_device(ctx, ctx.to_device(points), ctx.to_device(faces))
    torch.cuda.synchronize()
    assert dd == hd
    assert np.array_equal(dn.cpu().numpy(), hn)
    assert np.array_equal(db.cpu().numpy(), hb)
