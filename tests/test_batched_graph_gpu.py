"""Batched form prepared once and replayed as one HIP graph per batch (vgpu_batched_plan_*).

Every batch is checked two ways: per range against the oracle (the ranges' rows gathered into contiguous row
groups and run through vo_batched_filter_group_by_mt), and against the eager batched path over the same row
ranges (vgpu_batched_ranges_filter_group_by), which it must equal bit for bit, doubles included (same
compaction order, same stable sort, same reduction tiles). Graph replays are interleaved with eager work,
with synchronizations, and with in-place updates of the table, and each replay is checked against a fresh
eager result. Bit-exact for keys, COUNT, HUGEINT SUM, MIN, MAX; DOUBLE SUM/AVG within 1e-12 of the oracle."""
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
NAMES = ["count", "sum_v1", "min_v1", "max_v1", "sum_v2", "avg_v2"]

ROWS = 3_000_000
MAX_RANGES = 64
MAX_ROWS = 2_000_000


@pytest.fixture(scope="module")
def table(dev):
    import torch
    rng = np.random.default_rng(11)
    a = rng.integers(0, 10 ** 6, ROWS).astype(np.int32)
    c = rng.random(ROWS)
    # regions with different key spreads: a few groups, many groups, negative keys
    key = np.concatenate([rng.integers(0, 40, ROWS // 3), rng.integers(-5000, 200_000, ROWS // 3),
                          rng.integers(-(1 << 40), -(1 << 40) + 3000, ROWS - 2 * (ROWS // 3))]).astype(np.int64)
    v1 = rng.integers(-10 ** 6, 10 ** 6 + 1, ROWS).astype(np.int64)
    v2 = rng.standard_normal(ROWS)
    host = [a, c, key, v1, v2]
    cols = [torch.from_numpy(x).to(dev) for x in host]
    return host, cols


@pytest.fixture(scope="module")
def plan(dev, table):
    _, cols = table
    p = capi.BatchedPlan(cols, MAX_RANGES, MAX_ROWS, PREDS, 2, [3, 4], AGGS)
    yield p
    p.close()


def host_result(res):
    seg, keys, aggs, goffs = res
    out = {"seg": seg.cpu().numpy(), "key": keys.cpu().numpy(), "goffs": list(goffs)}
    for name, t in zip(NAMES, aggs):
        v = t.cpu().numpy()
        out[name] = v.reshape(-1, 2) if name == "sum_v1" else v
    return out


def eager(cols, starts, lens):
    return host_result(capi.batched_ranges_filter_group_by(cols, starts, lens, PREDS, 2, [3, 4], AGGS))


def check_oracle(host, starts, lens, got):
    """Per range against the oracle over the same rows laid out as contiguous row groups."""
    idx = np.concatenate([np.arange(s, s + n) for s, n in zip(starts, lens)] + [np.zeros(0, np.int64)])
    offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ref, goffs, _ = oracle.batched_filter_group_by_mt(*[x[idx] for x in host], offsets, 4)
    assert list(got["goffs"]) == [int(x) for x in goffs]
    g = len(ref["key"])
    assert len(got["key"]) == g
    np.testing.assert_array_equal(got["seg"], np.repeat(np.arange(len(starts)), np.diff(goffs)))
    np.testing.assert_array_equal(got["key"], ref["key"])
    for name in ("count", "min_v1", "max_v1", "sum_v1"):
        np.testing.assert_array_equal(got[name], ref[name], err_msg=name)
    for name in ("sum_v2", "avg_v2"):
        np.testing.assert_allclose(got[name], ref[name], rtol=1e-12, atol=0, err_msg=name)


def check_equal(got, want):
    for k in ("seg", "key", *NAMES):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    assert list(got["goffs"]) == list(want["goffs"])


def ranges(seed, nq, total):
    """nq ranges of log-uniform lengths holding about `total` rows, some empty, one row and one tile."""
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(0, np.log(max(2, 2 * total // max(nq, 1))), nq)).astype(np.int64)
    if lens.sum() > total:
        lens = lens * total // lens.sum()
    if nq > 3:
        lens[0], lens[1], lens[2] = 0, 1, 2048
    starts = rng.integers(0, ROWS - lens + 1)
    return [int(x) for x in starts], [int(x) for x in lens]


def test_plan_shape(plan):
    info = plan.info()
    assert info["kernel_nodes"] >= 8, info       # layout, state, filter, sort passes, reduce, offsets
    assert info["copy_nodes"] == 2, info         # descriptors in, results out
    assert info["other_nodes"] == 0, info        # no memset or host nodes
    assert info["group_bound"] <= MAX_ROWS


@pytest.mark.parametrize("seed,nq", [(1, 40), (2, 64), (3, 5), (4, 1)])
def test_graph_batch_matches_oracle_and_eager(table, plan, seed, nq):
    host, cols = table
    starts, lens = ranges(seed, nq, MAX_ROWS - 1000)
    got = host_result(plan.run(starts, lens))
    check_oracle(host, starts, lens, got)
    check_equal(got, eager(cols, starts, lens))


def test_graph_replays_interleaved_with_eager_work(dev, table, plan):
    import torch
    host, cols = table
    batches = [ranges(10 + i, nq, tot) for i, (nq, tot) in enumerate([(64, 1_500_000), (3, 10_000), (17, 400_000)])]
    # overlapping and repeated ranges in one batch
    batches.append(([0, 100, 100, ROWS - 4096, 7], [5000, 5000, 5000, 4096, 1]))
    for rep in range(2):
        for starts, lens in batches:
            got = host_result(plan.run(starts, lens))
            junk = torch.randn(1 << 20, device=dev).sum()          # eager work between replays
            torch.cuda.synchronize()
            want = eager(cols, starts, lens)
            check_equal(got, want)
            del junk
    check_oracle(host, *batches[-1], host_result(plan.run(*batches[-1])))


def test_graph_sees_table_updates(dev, table, plan):
    """The graph reads the table at every replay: an in-place update shows in the next run."""
    host, cols = table
    starts, lens = ranges(21, 24, 600_000)
    before = host_result(plan.run(starts, lens))
    cols[3].add_(7)
    try:
        after = host_result(plan.run(starts, lens))
        check_equal(after, eager(cols, starts, lens))
        np.testing.assert_array_equal(after["min_v1"], before["min_v1"] + 7)
        np.testing.assert_array_equal(after["count"], before["count"])
    finally:
        cols[3].sub_(7)
    check_equal(host_result(plan.run(starts, lens)), before)


def test_graph_empty_batches(table, plan):
    got = plan.run([], [])
    assert len(got[1]) == 0 and got[3] == [0]
    got = plan.run([5, 9, 0], [0, 0, 0])
    assert len(got[1]) == 0 and got[3] == [0, 0, 0, 0]


def test_graph_rejects_out_of_bound_batches_then_still_runs(table, plan):
    host, cols = table
    with pytest.raises(capi.VgpuError, match="exceed"):
        plan.run([0] * (MAX_RANGES + 1), [1] * (MAX_RANGES + 1))
    with pytest.raises(capi.VgpuError, match="bound"):
        plan.run([0, 0], [MAX_ROWS, 1])
    with pytest.raises(capi.VgpuError, match="past the table"):
        plan.run([ROWS - 10], [11])
    starts, lens = ranges(31, 9, 300_000)
    check_equal(host_result(plan.run(starts, lens)), eager(cols, starts, lens))


def test_plan_rejects_small_capacity(table):
    _, cols = table
    with pytest.raises(capi.VgpuError, match="capacity"):
        capi.BatchedPlan(cols, 8, 1000, PREDS, 2, [3, 4], AGGS, capacity=10)


def test_two_plans_and_count_only(dev, table):
    """A second plan (COUNT only: the run-length reduce) next to the first, on a 16-bit key."""
    import torch
    host, cols = table
    k16 = torch.from_numpy((host[2] % 1000).astype(np.int16)).to(dev)
    cols2 = [cols[0], cols[1], k16]
    aggs = [(capi.COUNT_STAR, -1, capi.INT64)]
    with capi.BatchedPlan(cols2, 16, 500_000, PREDS, 2, [], aggs) as p2:
        starts, lens = ranges(41, 16, 400_000)
        got = p2.run(starts, lens)
        want = capi.batched_ranges_filter_group_by(cols2, starts, lens, PREDS, 2, [], aggs)
        assert got[3] == want[3]
        assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1]) and torch.equal(got[2][0], want[2][0])
        keep = (host[0] < 500000) & (host[1] > 0.25)
        for q, (s, n) in enumerate(zip(starts, lens)):
            ks = (host[2][s:s + n] % 1000)[keep[s:s + n]]
            u, cnt = np.unique(ks, return_counts=True)
            lo, hi = got[3][q], got[3][q + 1]
            np.testing.assert_array_equal(got[1][lo:hi].cpu().numpy(), u.astype(np.int16))
            np.testing.assert_array_equal(got[2][0][lo:hi].cpu().numpy(), cnt)
