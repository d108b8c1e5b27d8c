"""vgpu_filter_gt_i32_noalloc (FilterGreaterThanInt32 with its scratch in a fixed per-device array) against
vgpu_filter_gt_i32 through the C ABI: bit-exact outputs and equal counts on calls in a row, the second over a
smaller input (the array keeps the first call's contents, so anything left over would show), and no growth of
the library's device pool. Oversized inputs are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from vectisdb_amd import capi

pytestmark = pytest.mark.gpu


def _gen(dev, n, seed):
    import torch
    x = torch.empty(n, dtype=torch.int32, device=dev)
    capi.generate_column(x, capi.INT32, 0, lo=-(1 << 30), hi=(1 << 30), seed=seed, stream_id=1)
    return x


@pytest.mark.parametrize("sizes", [(1 << 26, 3_000_001), (5_000_000, 77_777, 1, 4_194_304)])
def test_noalloc_matches_allocating_entry_point(dev, sizes):
    import torch
    out = torch.empty(max(sizes), dtype=torch.int32, device=dev)
    for i, n in enumerate(sizes):
        x = _gen(dev, n, 40 + i)
        threshold = 123_456 - 1000 * i
        ref = capi.filter_gt_i32(x, threshold)
        torch.cuda.synchronize()
        before = capi.pool_stats()
        out.fill_(-7)
        kept = capi.filter_gt_i32_noalloc(x, out, threshold)
        after = capi.pool_stats()
        assert kept == ref.numel()
        assert torch.equal(out[:kept], ref)
        # nothing written past the kept rows
        assert bool((out[kept:] == -7).all())
        assert after["in_use"] == before["in_use"] and after["cached"] == before["cached"]
        # independent check of the kept set on the host
        hx = x.cpu().numpy()
        np.testing.assert_array_equal(out[:kept].cpu().numpy(), hx[hx > threshold])


def test_noalloc_empty_and_extreme_thresholds(dev):
    import torch
    x = _gen(dev, 100_000, 9)
    out = torch.empty(100_000, dtype=torch.int32, device=dev)
    assert capi.filter_gt_i32_noalloc(x[:0], out, 0) == 0
    assert capi.filter_gt_i32_noalloc(x, out, 2**31 - 1) == 0
    assert capi.filter_gt_i32_noalloc(x, out, -2**31) == int((x.cpu().numpy() > -2**31).sum())
    assert torch.equal(out[:100_000][x > -2**31], x[x > -2**31])


def test_noalloc_rejects_inputs_beyond_its_scratch(dev):
    import torch
    lib = capi.load()
    limit = lib.vgpu_filter_gt_i32_noalloc_max_rows()
    assert limit == 1 << 31
    x = torch.zeros(16, dtype=torch.int32, device=dev)
    count = C.c_uint64(99)
    # the size check runs before any pointer is used or any work is queued
    st = lib.vgpu_filter_gt_i32_noalloc(C.c_void_p(x.data_ptr()), limit + 1, 0, C.c_void_p(x.data_ptr()),
                                        C.byref(count), None)
    assert st != 0
    assert b"fixed scratch limit" in lib.vgpu_last_error()
    assert count.value == 99
