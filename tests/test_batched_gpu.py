"""Batched form: many filter + GROUP BY queries over the row groups of one table in one pipeline
(vgpu_batched_filter_group_by), checked per query against the oracle AND against the per-query path
(vgpu_filter_project then vgpu_group_by on that row group alone): each query's result must be the same
alone and in the batch. Bit-exact for keys, COUNT, HUGEINT SUM, MIN, MAX; DOUBLE SUM/AVG within 1e-12."""
import numpy as np
import pytest

import oracle
from vectisdb_amd import capi

pytestmark = pytest.mark.gpu

AGGS = [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT), (capi.MIN, 0, capi.INT64),
        (capi.MAX, 0, capi.INT64), (capi.SUM, 1, capi.FLOAT64), (capi.AVG, 1, capi.FLOAT64)]
# columns: 0 a INT32, 1 c DOUBLE, 2 key INT64, 3 v1 INT64, 4 v2 DOUBLE
PREDS = [("((bool*)out)[idx] = (((const int*)inputs[0])[row]) < 500000;", [0], capi.BOOLEAN),
         ("((bool*)out)[idx] = (((const double*)inputs[0])[row]) > 0.25;", [1], capi.BOOLEAN)]


def hugeints(t):
    a = t.cpu().numpy().reshape(-1, 2)
    return [int(lo) % (1 << 64) + (int(hi) << 64) for lo, hi in a]


def make_table(rng, sizes, max_groups):
    n = int(sum(sizes))
    a = rng.integers(0, 10 ** 6, n).astype(np.int32)
    c = rng.random(n)
    key = np.concatenate([rng.integers(0, max(1, int(g)), int(s)) for s, g in zip(sizes, max_groups)]
                         ).astype(np.int64) if n else np.zeros(0, np.int64)
    v1 = rng.integers(-10 ** 6, 10 ** 6 + 1, n).astype(np.int64)
    v2 = rng.standard_normal(n)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return a, c, key, v1, v2, offsets


def run_batched(dev, a, c, key, v1, v2, offsets):
    import torch
    cols = [torch.from_numpy(x).to(dev) for x in (a, c, key, v1, v2)]
    seg, keys, aggs, goffs = capi.batched_filter_group_by(cols, torch.from_numpy(offsets).to(dev), PREDS, 2, [3, 4],
                                                          AGGS)
    return seg.cpu().numpy(), keys.cpu().numpy(), aggs, goffs


def per_query(dev, a, c, key, v1, v2):
    """The plain path for one query: fused filter/projection, then GROUP BY, through the C ABI."""
    import torch
    if len(a) == 0:
        return np.zeros(0, np.int64), None, 0
    cols = [torch.from_numpy(x).to(dev) for x in (a, c, key, v1, v2)]
    projs = [("((long long*)out)[idx] = ((const long long*)inputs[0])[row];", [2], capi.INT64),
             ("((long long*)out)[idx] = ((const long long*)inputs[0])[row];", [3], capi.INT64),
             ("((double*)out)[idx] = ((const double*)inputs[0])[row];", [4], capi.FLOAT64)]
    outs, _, _, kept = capi.filter_project(cols, PREDS, projs)
    if kept == 0:
        return np.zeros(0, np.int64), None, 0
    keys, aggs, g = capi.group_by([outs[0][:kept]], [outs[1][:kept], outs[2][:kept]], AGGS, kept)
    return keys[0].cpu().numpy(), aggs, g


def check_against_oracle(keys, count, s1, mn, mx, s2, avg, a, c, key, v1, v2):
    keep = (a < 500000) & (c > 0.25)
    if not keep.any():
        assert len(keys) == 0
        return
    ref = oracle.group_by_i64(key[keep], v1[keep], v2[keep])
    np.testing.assert_array_equal(keys, ref["key"])
    np.testing.assert_array_equal(count, ref["count"])
    assert s1 == ref["sum_v1"]
    np.testing.assert_array_equal(mn, ref["min_v1"])
    np.testing.assert_array_equal(mx, ref["max_v1"])
    np.testing.assert_allclose(s2, ref["sum_v2"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(avg, ref["avg_v2"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("seed,nq", [(0, 40), (1, 7), (2, 1)])
def test_batched_matches_oracle_and_per_query_path(dev, seed, nq):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1), np.log(200_000), nq)).astype(np.int64)
    if nq > 3:
        sizes[1] = 0       # an empty row group
        sizes[2] = 1       # a single row
        sizes[3] = 2048    # exactly one tile
    groups = np.exp(rng.uniform(np.log(16), np.log(50_000), nq)).astype(np.int64)
    a, c, key, v1, v2, offsets = make_table(rng, sizes, groups)
    seg, keys, aggs, goffs = run_batched(dev, a, c, key, v1, v2, offsets)
    assert len(goffs) == nq + 1 and goffs[0] == 0 and goffs[-1] == len(keys)
    assert (np.diff(seg) >= 0).all()
    count = aggs[0].cpu().numpy()
    s1 = hugeints(aggs[1])
    mn, mx = aggs[2].cpu().numpy(), aggs[3].cpu().numpy()
    s2, avg = aggs[4].cpu().numpy(), aggs[5].cpu().numpy()
    for q in range(nq):
        lo, hi = goffs[q], goffs[q + 1]
        assert (seg[lo:hi] == q).all()
        r0, r1 = offsets[q], offsets[q + 1]
        check_against_oracle(keys[lo:hi], count[lo:hi], s1[lo:hi], mn[lo:hi], mx[lo:hi], s2[lo:hi], avg[lo:hi],
                             a[r0:r1], c[r0:r1], key[r0:r1], v1[r0:r1], v2[r0:r1])
        # Same result alone (per-query path) as in the batch: bit-exact, doubles within 1e-12.
        pk, paggs, pg = per_query(dev, a[r0:r1], c[r0:r1], key[r0:r1], v1[r0:r1], v2[r0:r1])
        assert pg == hi - lo
        if pg:
            np.testing.assert_array_equal(pk, keys[lo:hi])
            np.testing.assert_array_equal(paggs[0].cpu().numpy(), count[lo:hi])
            assert hugeints(paggs[1]) == s1[lo:hi]
            np.testing.assert_array_equal(paggs[2].cpu().numpy(), mn[lo:hi])
            np.testing.assert_array_equal(paggs[3].cpu().numpy(), mx[lo:hi])
            np.testing.assert_allclose(paggs[4].cpu().numpy(), s2[lo:hi], rtol=1e-12, atol=0)
            np.testing.assert_allclose(paggs[5].cpu().numpy(), avg[lo:hi], rtol=1e-12, atol=0)


def test_batched_rejects_bad_arguments(dev):
    import torch
    x = torch.zeros(10, dtype=torch.int64, device=dev)
    offs = torch.tensor([0, 10], dtype=torch.int64, device=dev)
    with pytest.raises(capi.VgpuError):
        capi.batched_filter_group_by([x, x], offs, [], 5, [1], AGGS[:1])   # key column out of range
    d = torch.zeros(10, dtype=torch.float64, device=dev)
    with pytest.raises(capi.VgpuError):
        capi.batched_filter_group_by([d, x], offs, [], 0, [1], AGGS[:1])   # float key


def test_native_per_query_loop_matches_batched(dev):
    """vgpu_per_query_filter_group_by (the per-query baseline the batched form is timed against) writes the
    same layout; results agree with the batched launch (bit-exact, doubles within 1e-12)."""
    import torch
    rng = np.random.default_rng(11)
    nq = 24
    sizes = np.exp(rng.uniform(np.log(1), np.log(100_000), nq)).astype(np.int64)
    sizes[5] = 0
    groups = np.exp(rng.uniform(np.log(16), np.log(20_000), nq)).astype(np.int64)
    a, c, key, v1, v2, offsets = make_table(rng, sizes, groups)
    cols = [torch.from_numpy(x).to(dev) for x in (a, c, key, v1, v2)]
    offs = torch.from_numpy(offsets).to(dev)
    b = capi.batched_filter_group_by(cols, offs, PREDS, 2, [3, 4], AGGS)
    p = capi.batched_filter_group_by(cols, offs, PREDS, 2, [3, 4], AGGS, per_query=True)
    assert b[3] == p[3]
    np.testing.assert_array_equal(b[0].cpu().numpy(), p[0].cpu().numpy())
    np.testing.assert_array_equal(b[1].cpu().numpy(), p[1].cpu().numpy())
    for i, (_, _, t) in enumerate(AGGS):
        x, y = b[2][i].cpu().numpy(), p[2][i].cpu().numpy()
        if t == capi.FLOAT64:
            np.testing.assert_allclose(y, x, rtol=1e-12, atol=0)
        else:
            np.testing.assert_array_equal(y, x)
