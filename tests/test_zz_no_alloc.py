"""mt_compact_noalloc (the compaction whose look-back workspace is a fixed
per-device array in the library) against mt_compact, called through the C
ABI: the same survivors, order, count and payload bytes on two calls in a
row, the second on a smaller input, with a payload of every element size;
and a capacity above MT_NOALLOC_MAX_ELEMS is refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from mambatracer_amd import _lib

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 28


def _compact(ctx, name, flags, length, payloads):
    n = int(flags.shape[0])
    idx = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    count = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    arr = (_lib.Payload * max(1, len(payloads)))()
    dsts = []
    for j, src in enumerate(payloads):
        dst = torch.zeros_like(src)
        eb = src.element_size() * int(np.prod(src.shape[1:], dtype=np.int64))
        arr[j] = _lib.Payload(src.data_ptr(), dst.data_ptr(), eb)
        dsts.append(dst)
    _lib.call(name, ctx.handle, flags.data_ptr(), 0, length.data_ptr(), n, idx.data_ptr(), count.data_ptr(),
              arr, len(payloads))
    return idx, count, dsts


@pytest.mark.parametrize("sizes", [(3_000_001, 70_000), (40_000, 8192)])
def test_noalloc_matches_compact_twice(ctx, sizes):
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    for call_no, n in enumerate(sizes):
        print(f"[no-alloc] call {call_no}: n={n}", flush=True)
        flags = (torch.rand(n, device="cuda", generator=gen) < 0.37).to(torch.int32)
        length = torch.tensor([n - 5], dtype=torch.int32, device="cuda")  # a live length below capacity
        p4 = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=gen)
        p8 = torch.randint(-2**31, 2**31 - 1, (n, 2), dtype=torch.int32, device="cuda", generator=gen)
        p16 = torch.rand((n, 4), device="cuda", generator=gen)
        ref = _compact(ctx, "mt_compact", flags, length, [p4, p8, p16])
        got = _compact(ctx, "mt_compact_noalloc", flags, length, [p4, p8, p16])
        torch.cuda.synchronize()
        ctx.check_scans()
        cnt = int(ref[1].item())
        live = flags[: n - 5].cpu().numpy()
        assert cnt == int(live.sum()) == int(got[1].item())
        assert torch.equal(ref[0][:cnt], got[0][:cnt])
        assert np.array_equal(got[0][:cnt].cpu().numpy(), np.flatnonzero(live))
        for a, b in zip(ref[2], got[2]):
            assert torch.equal(a[:cnt].view(torch.int32), b[:cnt].view(torch.int32))


def test_noalloc_rejects_capacity_above_maximum(ctx):
    flags = torch.zeros(4, dtype=torch.int32, device="cuda")
    length = torch.tensor([4], dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = _lib.load().mt_compact_noalloc(ctx.handle, flags.data_ptr(), 0, length.data_ptr(), NOALLOC_MAX + 1,
                                        None, count.data_ptr(), None, 0)
    assert rc == -1  # MT_ERR_ARG
    assert "exceeds the maximum" in _lib.last_error()
