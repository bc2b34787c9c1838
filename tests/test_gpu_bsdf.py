"""Stage B on the GPU against the oracle: setupBsdf, SampleBsdf and
EvaluateBsdf.  Bar: bit-exact f, pdf, cos and directions (both sides run the
same IEEE operations, the transcendental functions included: mt_xmath.h);
NaN positions equal.  Any mismatch is listed."""
import numpy as np
import pytest
import torch

from mambatracer_amd import api
from oracle import orc_py as orc

pytestmark = pytest.mark.gpu


def materials(n, seed):
    """Config-1 materials: albedo U[0.2,0.9]^3, F0 0.04, alpha U[0.05,1] (some anisotropic)."""
    rng = np.random.default_rng(seed)
    m = np.zeros((n, 36), np.float32)
    m[:, 4:7] = rng.uniform(0.2, 0.9, size=(n, 3))
    m[:, 8] = 1.0                      # diffuse_param.x
    m[:, 12:15] = 0.04                 # glossy_color = F0
    m[:, 16] = 1.0                     # glossy_param.x
    a = rng.uniform(0.05, 1.0, size=n)
    m[:, 17] = a
    m[:, 18] = np.where(rng.random(n) < 0.3, rng.uniform(0.05, 1.0, size=n), a)
    return m


def unit(rng, n):
    v = rng.normal(size=(n, 4))
    v[:, 3] = 0
    v /= np.linalg.norm(v[:, :3], axis=1, keepdims=True)
    return v.astype(np.float32)


def bit_mismatches(name, a, b):
    """Entries whose bits differ (NaN matches NaN: the device's and the host's
    default NaN differ in sign); prints the first few."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    bad = (a.view(np.int32) != b.view(np.int32)) & ~(np.isnan(a) & np.isnan(b))
    if bad.any():
        idx = np.argwhere(bad)[:6]
        print(f"{name}: {int(bad.sum())} differ, e.g. {idx.tolist()} {a[tuple(idx.T)].tolist()} vs "
              f"{b[tuple(idx.T)].tolist()}", flush=True)
    return int(bad.sum())


def _setup(ctx, n=100000, seed=0):
    rng = np.random.default_rng(seed)
    M = materials(64, seed)
    nrm = unit(rng, n)
    ray_d = unit(rng, n)
    ray_d[:, :3] = np.where((np.sum(ray_d[:, :3] * nrm[:, :3], 1) > 0)[:, None], -ray_d[:, :3], ray_d[:, :3])
    ng = nrm.copy()
    mat = rng.integers(0, 64, size=n).astype(np.int32)
    mat[::97] = -1                 # misses: no material is read, mat_id -1
    rec, mid = api.bsdf_setup(ctx, ctx.to_device(ray_d), ctx.to_device(nrm), ctx.to_device(ng), ctx.to_device(mat),
                              ctx.to_device(M))
    orec, omid = orc.bsdf_setup(ray_d, nrm, ng, mat, M)
    return rng, M, rec, mid, orec, omid


def test_bsdf_setup(ctx):
    _, _, rec, mid, orec, omid = _setup(ctx)
    np.testing.assert_array_equal(mid.cpu().numpy(), omid)
    np.testing.assert_array_equal(rec.cpu().numpy().view(np.int32), orec.view(np.int32))


def test_bsdf_sample(ctx):
    print("[case] bsdf sample", flush=True)
    rng, M, rec, mid, orec, omid = _setup(ctx, seed=1)
    n = rec.shape[0]
    u = rng.random((n, 3)).astype(np.float32)
    f, d, pdf, cs = api.bsdf_sample(ctx, rec, mid, ctx.to_device(M), ctx.to_device(u))
    of, od, opdf, ocs = orc.bsdf_sample(orec, omid, M, u)
    f, d, pdf, cs = (x.cpu().numpy() for x in (f, d, pdf, cs))
    for name, a, b in [("f", f, of), ("dir", d, od), ("pdf", pdf, opdf), ("cos", cs, ocs)]:
        assert bit_mismatches(name, a, b) == 0


def test_bsdf_eval(ctx):
    print("[case] bsdf eval", flush=True)
    rng, M, rec, mid, orec, omid = _setup(ctx, seed=2)
    n = rec.shape[0]
    dirs = unit(rng, n)[:, :3].copy()
    f, pdf, cs = api.bsdf_eval(ctx, rec, mid, ctx.to_device(M), ctx.to_device(dirs))
    of, opdf, ocs = orc.bsdf_eval(orec, omid, M, dirs)
    for name, a, b in [("f", f.cpu().numpy(), of), ("pdf", pdf.cpu().numpy(), opdf), ("cos", cs.cpu().numpy(), ocs)]:
        assert bit_mismatches(name, a, b) == 0
