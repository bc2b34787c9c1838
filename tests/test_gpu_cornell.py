"""The reference's own default scene, the Cornell box of
TracerAPI.initDefaultMesh (TracerAPI.java:205-371; fixture written by
tests/golden/make_cornell.py): OBJ text -> mt_obj_parse -> BVH -> mesh-light
list -> bounce loop on the GPU, against the oracle's loop on the same arrays.

Bar: the parsed mesh has the fixture's vertices, normals, triangle count and
one material per `o` group; the emitter list is exactly the light's
triangles; images and ray counts are bit-exact against the oracle for both
light-pick rules (the reference's lightInfo[0], Render.cl:305, and the
picked light); the red wall (leftWall, material 7) lights the left of the
frame and the green wall (rightWall, material 3) the right."""
import numpy as np
import pytest
import torch

from mambatracer_amd import _lib, api
from mambatracer_amd import workloads as W
from oracle import orc_py as orc
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def cornell_obj(fx):
    lines = [f"v {x} {y} {z}" for x, y, z in fx["v"]] + [f"vn {x} {y} {z}" for x, y, z in fx["vn"]]
    group = -1
    for p in fx["polygons"]:
        if p["group"] != group:
            group = p["group"]
            lines.append(f"o {fx['groups'][group]}")
        lines.append("f " + " ".join(f"{v}//{n}" for v, n in p["corners"]))
    return "\n".join(lines) + "\n"


def cornell_materials(fx):
    d = fx["defaults"]
    mats = np.zeros((len(fx["groups"]), 36), np.float32)
    for g in range(len(fx["groups"])):
        m = mats[g]
        m[4:7] = d["diffuse"]
        m[8] = d["diffuse_param_x"]
        m[12:15] = d["glossy"]
        m[16:19] = d["glossy_param"]
        m[20:23] = d["mirror"]
        m[28:31] = d["emission"]
        m[32:35] = d["emission_param"]
        spec = fx["materials"].get(str(g), {})
        if "diffuse" in spec:
            m[4:7] = spec["diffuse"]
        if "emitter" in spec:
            m[28:31] = spec["emitter"]
            m[32:35] = d["set_emitter_param"]
    return mats


@pytest.fixture(scope="module")
def cornell():
    fx = golden("cornell_box.json")
    points, normals, uvs, faces, ngroups = api.OBJParser().readString(cornell_obj(fx))
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    return fx, points, normals, faces, nodes, bounds, depth, ngroups


def test_cornell_mesh_parses_as_the_fixture(cornell):
    fx, points, normals, faces, nodes, bounds, depth, ngroups = cornell
    assert points.shape[0] == len(fx["v"]) and normals.shape[0] == len(fx["vn"])
    np.testing.assert_array_equal(points[:, :3], np.asarray(fx["v"], np.float32))
    ntri = sum(len(p["corners"]) - 2 for p in fx["polygons"])
    assert faces.shape[0] == ntri and ngroups == len(fx["groups"])
    # material of each triangle = its polygon's group
    want = np.concatenate([[p["group"]] * (len(p["corners"]) - 2) for p in fx["polygons"]])
    np.testing.assert_array_equal(faces[:, 9] & 0xFFFF, want)


@pytest.mark.parametrize("flags", [0, _lib.MT_RENDER_PICKED_LIGHT])
def test_cornell_render_matches_oracle(ctx, cornell, flags):
    print(f"[case] cornell box flags={flags}", flush=True)
    fx, points, normals, faces, nodes, bounds, depth, _ = cornell
    mats = cornell_materials(fx)
    sc = api.Scene(ctx, points, faces, nodes, bounds, depth, normals=normals)
    dm = ctx.to_device(mats)
    lf, nl = api.mesh_lights(ctx, sc, dm)
    torch.cuda.synchronize()
    n = int(nl.item())
    light_tris = np.flatnonzero((faces[:, 9] & 0xFFFF) == 6)
    assert n == len(light_tris) and np.array_equal(lf.cpu().numpy()[:n], light_tris)
    job = W.RenderJob(width=96, height=96, spp=8, max_bounces=6, rr_start=3, seed=3, pos=(0.0, 0.0, -3.4),
                      look=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), vfov_deg=45.0, flags=flags)
    img, st = api.render(ctx, sc, dm, None, job, light_faces=lf, num_lights=nl)
    img = img.cpu().numpy()[:, :3]
    pids = np.arange(job.width * job.height * job.spp, dtype=np.int32)
    L, rays = orc.render_paths(points, normals, faces, nodes, bounds, mats, None, job, pids, light_faces=light_tris)
    oimg = orc.resolve(L, job.width, job.height, job.spp)
    bad = np.any(img.view(np.int32) != oimg.view(np.int32), axis=1)
    print(f"rays gpu {st.extension_rays}+{st.shadow_rays} oracle {rays[0]}+{rays[1]}, pixels differing "
          f"{int(bad.sum())}, mean {img.mean(0)}", flush=True)
    assert (int(st.extension_rays), int(st.shadow_rays)) == (int(rays[0]), int(rays[1]))
    assert not bad.any()
    assert np.isfinite(img).all() and img.mean() > 0
    im = img.reshape(job.height, job.width, 3)
    left, right = im[:, : job.width // 4].mean((0, 1)), im[:, -job.width // 4:].mean((0, 1))
    assert left[0] > left[1] and right[1] > right[0], (left, right)
