"""Stage D on the GPU: single-pass exclusive scan and stable compaction
against the oracle, including the reference's all-ones known answer
(test/algorithm/PrefixSumIntegerTest.java:30-36) and edge sizes."""
import numpy as np
import pytest
import torch

from mambatracer_amd import api
from oracle import orc_py as orc
from tests.conftest import golden

pytestmark = pytest.mark.gpu

# tiles are 8192 elements: sizes straddle one, two and many tiles
SIZES = [0, 1, 2, 63, 64, 65, 200, 2047, 2048, 2049, 4096 * 3 + 17, 8191, 8192, 8193, 8192 * 3 + 17, 1 << 20,
         3_000_001]


@pytest.mark.parametrize("n", SIZES)
def test_scan_random(ctx, n):
    print(f"[case] scan n={n}", flush=True)
    rng = np.random.default_rng(n)
    x = rng.integers(-1000, 1000, size=n).astype(np.int32)
    ps = api.PrefixSumInteger(ctx, ctx.to_device(x))
    ps.execute()
    total = ps.getCount()
    ref, rtot = orc.scan_excl(x)
    np.testing.assert_array_equal(ps.getPrefixSum().cpu().numpy(), ref)
    assert total == rtot


def test_scan_all_ones_known_answer(ctx):
    """PrefixSumIntegerTest: 200 ones -> out[i] = i, total = 200."""
    k = golden("prefix_sum_all_ones.json")
    ps = api.PrefixSumInteger(ctx, ctx.to_device(np.array(k["input"], np.int32)))
    ps.execute()
    assert ps.getCount() == k["total"]
    np.testing.assert_array_equal(ps.getPrefixSum().cpu().numpy(), k["exclusive_scan"])


def test_scan_device_length_shorter_than_capacity(ctx):
    x = np.arange(10000, dtype=np.int32) % 7
    dx = ctx.to_device(x)
    out = torch.full((10000,), -5, dtype=torch.int32, device=ctx.torch_device)
    ln = ctx.scalar_i32(4321)
    tot = ctx.scalar_i32(-1)
    api.call("mt_scan_excl_i32", ctx.handle, dx.data_ptr(), out.data_ptr(), ln.data_ptr(), 10000, tot.data_ptr())
    o = out.cpu().numpy()
    ref, rtot = orc.scan_excl(x[:4321])
    np.testing.assert_array_equal(o[:4321], ref)
    assert (o[4321:] == -5).all()
    assert int(tot.item()) == rtot


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("density", [0.0, 0.3, 1.0])
def test_compact_stable(ctx, n, density):
    print(f"[case] compact n={n} p={density}", flush=True)
    rng = np.random.default_rng(n + int(density * 10))
    pred = (rng.random(n) < density).astype(np.int32)
    pay4 = rng.integers(0, 1 << 30, size=n).astype(np.int32)
    pay16 = rng.random((n, 4)).astype(np.float32)
    idx, cnt, (d4, d16) = api.compact(ctx, ctx.to_device(pred), payloads=[ctx.to_device(pay4), ctx.to_device(pay16)])
    oidx, ocnt = orc.compact_stable(pred)
    assert int(cnt.item()) == ocnt
    np.testing.assert_array_equal(idx.cpu().numpy()[:ocnt], oidx)
    np.testing.assert_array_equal(d4.cpu().numpy()[:ocnt], pay4[oidx])
    np.testing.assert_array_equal(d16.cpu().numpy()[:ocnt], pay16[oidx])
    ctx.check_scans()


def test_compact_prim_mode(ctx):
    """flag_mode 1 keeps slots whose value is >= 0 (a hit's primitive id)."""
    rng = np.random.default_rng(11)
    prim = rng.integers(-1, 5, size=100000).astype(np.int32)
    idx, cnt, _ = api.compact(ctx, ctx.to_device(prim), flag_mode=1)
    oidx, ocnt = orc.compact_stable((prim >= 0).astype(np.int32))
    assert int(cnt.item()) == ocnt
    np.testing.assert_array_equal(idx.cpu().numpy()[:ocnt], oidx)


def test_ccompact_mirror(ctx):
    """CCompact.execute: hit slots' pixel ids packed to the front, count set."""
    rng = np.random.default_rng(3)
    n = 250000
    hit = (rng.random(n) < 0.6).astype(np.int32)
    pixels = rng.permutation(n).astype(np.int32)
    dh, dp = ctx.to_device(hit), ctx.to_device(pixels)
    cnt = ctx.scalar_i32(n)
    cc = api.CCompact()
    cc.init(ctx, dh, dp, cnt)
    cc.execute()
    oidx, ocnt = orc.compact_stable(hit)
    assert int(cnt.item()) == ocnt
    np.testing.assert_array_equal(dp.cpu().numpy()[:ocnt], pixels[oidx])


@pytest.mark.parametrize("n,length", [(250000, 250000), (100000, 73001), (1024, 1024), (5000, 0)])
def test_ccompact_records(ctx, n, length):
    """CCompact.execute on 96-byte Intersection records in place
    (mt_compact_isects): hit records and their pixel ids packed to the front
    in slot order (the oracle's stable compaction), the rest of the first
    `length` records re-initialised (InitIsect, pixel 0), later slots
    untouched, count set; run twice on the same buffers (temp reuse)."""
    print(f"[case] ccompact records n={n} length={length}", flush=True)
    rng = np.random.default_rng(7 + n)
    rec = rng.integers(-2**31, 2**31 - 1, size=(n, 24), dtype=np.int64).astype(np.int32)
    rec[:, 20] = np.where(rng.random(n) < 0.55, rng.integers(1, 5, n), 0)   # hit markers, nonzero = hit
    pixels = rng.permutation(n).astype(np.int32)
    di, dp = ctx.to_device(rec), ctx.to_device(pixels)
    cnt = ctx.scalar_i32(-9)
    cc = api.CCompact()
    cc.init(ctx, di, dp, cnt)
    cc.length.fill_(length)
    for rep in range(2):
        if rep == 1:   # second pass: the survivors are all hits, the tail misses
            rec, pixels = di.cpu().numpy().copy(), dp.cpu().numpy().copy()
        cc.execute()
        got_r, got_p = di.cpu().numpy(), dp.cpu().numpy()
        oidx, ocnt = orc.compact_stable((rec[:length, 20] != 0).astype(np.int32))
        assert int(cnt.item()) == ocnt
        want = rec[oidx].copy()
        want[:, 21:24] = 0
        np.testing.assert_array_equal(got_r[:ocnt], want)
        np.testing.assert_array_equal(got_p[:ocnt], pixels[oidx])
        tail = np.zeros((length - ocnt, 24), np.int32)
        tail[:, 18] = -1
        tail[:, 19] = -1
        np.testing.assert_array_equal(got_r[ocnt:length], tail)
        np.testing.assert_array_equal(got_p[ocnt:length], 0)
        np.testing.assert_array_equal(got_r[length:], rec[length:])
        np.testing.assert_array_equal(got_p[length:], pixels[length:])


def test_mesh_lights(ctx):
    """identifyMeshLights + prepareLightInfo: emitter faces in index order."""
    rp, rf = api.gen_room(0)
    rf[[1, 5, 6], 9] = 1
    nodes, bounds, depth = api.CNormalBVH.build(rp, rf)
    sc = api.Scene(ctx, rp, rf, nodes, bounds, depth)
    mats = np.zeros((2, 36), np.float32)
    mats[1, 28:31] = 1.0
    mats[1, 32] = 1.0  # emission_param.x
    out, cnt = api.mesh_lights(ctx, sc, ctx.to_device(mats))
    assert int(cnt.item()) == 3
    np.testing.assert_array_equal(out.cpu().numpy()[:3], [1, 5, 6])
