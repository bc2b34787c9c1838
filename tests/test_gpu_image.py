"""Image pipeline (mt_image_process) and float scan (mt_scan_excl_f32) on the
GPU against the oracle.

Bar: the float scan and its total bit-exact (same summation order, IEEE
adds); log-luminance count exact, total within 1e-5 relative and frame
values within 1e-5 relative (device logf/expf/powf may differ from the host
libm by an ulp); ARGB bytes equal except where such an ulp crosses a
1/255 step (at most 0.1% of pixels, and then by one)."""
import numpy as np
import pytest
import torch

from mambatracer_amd import api
from mambatracer_amd import workloads as W
from oracle import orc_py as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 777, 2048, 2049, 70_000, 1_000_003])
def test_float_scan_bit_exact(ctx, n):
    print(f"[case] fscan n={n}", flush=True)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1.0, 3.0, size=n).astype(np.float32)
    ps = api.PrefixSumFloat(ctx, ctx.to_device(x))
    ps.execute()
    torch.cuda.synchronize()
    out, tot = orc.scan_excl_f32(x)
    np.testing.assert_array_equal(ps.getPrefixSum().cpu().numpy().view(np.int32), out.view(np.int32))
    assert float(ps.getCTotal().item()) == tot


def test_float_scan_device_length(ctx):
    """*d_len < capacity: only the live prefix is scanned and totalled."""
    x = np.arange(10_000, dtype=np.float32) * 0.25
    dx = ctx.to_device(x)
    out = torch.full((10_000,), -1.0, dtype=torch.float32, device=ctx.torch_device)
    tot = torch.zeros((1,), dtype=torch.float32, device=ctx.torch_device)
    dlen = ctx.scalar_i32(4321)
    api.call("mt_scan_excl_f32", ctx.handle, dx.data_ptr(), out.data_ptr(), dlen.data_ptr(), 10_000, tot.data_ptr())
    torch.cuda.synchronize()
    o, t = orc.scan_excl_f32(x[:4321])
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got[:4321], o)
    assert (got[4321:] == -1.0).all()
    assert float(tot.item()) == t


def _compare(frame, argb, tot, oframe, oargb, otot):
    assert tot[1] == otot[1]
    assert abs(tot[0] - otot[0]) <= 1e-5 * abs(otot[0])
    fin = np.isfinite(oframe)
    assert (np.isfinite(frame) == fin).all()
    np.testing.assert_allclose(frame[fin], oframe[fin], rtol=1e-5, atol=1e-7)
    diff = argb != oargb
    assert diff.mean() <= 1e-3, diff.mean()
    if diff.any():
        a = argb[diff].view(np.uint8).reshape(-1, 4).astype(int)
        b = oargb[diff].view(np.uint8).reshape(-1, 4).astype(int)
        assert np.abs(a - b).max() <= 1


def test_image_process_matches_oracle(ctx):
    """A rendered frame (c1 scene, reduced) through CImage.processImage, then
    a later frame that reuses the stored totals."""
    print("[case] image pipeline", flush=True)
    pts, nrm, fc, nodes, bounds, depth = W.c1_scene(count=24)
    sc = api.Scene(ctx, pts, fc, nodes, bounds, depth, normals=nrm)
    env = api.CEnvironment(ctx)
    env.setEnvironmentMap(W.env_map())
    job = W.RenderJob(width=200, height=120, spp=2, pos=(0.0, 0.5, -8.0), look=(0.0, -1.0, 0.0))
    img, _ = api.render(ctx, sc, ctx.to_device(W.c1_materials()), env, job, stats=False)
    im = api.CImage(ctx, job.width, job.height)
    im.cFrameAccum.copy_(img * 2.0)           # as if two frames had been accumulated
    im.frameCount = 2.0
    im.processImage()
    torch.cuda.synchronize()
    accum = (img * 2.0).cpu().numpy()
    oframe, oargb, otot = orc.image_process(accum, 2.0, 0.18, 2.2)
    tot = (float(im.cTotalLogLuminance.item()), int(im.cTotalNumber.item()))
    _compare(im.cFrameBuffer.cpu().numpy(), im.getFrameARGB().cpu().numpy(), tot, oframe, oargb, otot)
    # frame 3: totals reused, different exposure and gamma
    im.frameCount = 3.0
    im.setExposure(0.5)
    im.setGamma(1.8)
    im.processImage()
    torch.cuda.synchronize()
    oframe3, oargb3, otot3 = orc.image_process(accum, 3.0, 0.5, 1.8, argb=oargb, totals=tot)
    assert otot3 == tot
    _compare(im.cFrameBuffer.cpu().numpy(), im.getFrameARGB().cpu().numpy(), tot, oframe3, oargb3, otot3)


def test_image_init_and_black_frame(ctx):
    im = api.CImage(ctx, 64, 32)
    im.cFrameAccum.zero_()
    im.processImage()                         # no lit pixel: argb keeps its initial value
    torch.cuda.synchronize()
    assert (im.getFrameARGB().cpu().numpy() == orc.image_init_argb()).all()
    assert int(im.cTotalNumber.item()) == 0
