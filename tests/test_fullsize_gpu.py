"""Parity at the measured configurations' full sizes, on the same device-generated data the benchmark times.

* GROUP BY over 100M rows (Zipf(1.1) keys over 1M values, 7 aggregates): every group against the
  multithreaded oracle (vo_group_by_i64_mt, itself pinned bit for bit against the single-threaded restatement).
* Join + aggregate over 2^24 build x 2^28 probe rows: SUM/COUNT against the closed form of a permutation build
  side (per-key probe counts by bincount, the payload looked up through the inverse permutation). The build
  span of exactly 2^24 keys makes each of the 256 key-range buckets a 65,536-entry 16-bit table slice that
  fills the probe kernel's LDS; one build payload raised past 12 bits forces the 64-bit table through L2.
  Both forms are asserted by the profile name of the probe launch.
* Cross product + filter over 4,096 x 65,536 pairs: count and sums of the kept pairs against the oracle's
  binary-search count; the left-major order of every kept pair against its closed form.
* Batched form over 512 row groups of 2K..4M rows: every query against the oracle on its own rows.

Bars: bit-exact for integer work (keys, COUNT, HUGEINT SUM, MIN, MAX, row sets); DOUBLE SUM/AVG within 1e-12
relative of the correctly rounded sum.
"""
import os

import numpy as np
import pytest

import oracle
from vectisdb_amd import capi, workloads

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _hugeint_pairs(t, g):
    return t[: 2 * g].cpu().numpy().view(np.int64).reshape(-1, 2)


def test_groupby_100m_rows_every_group(dev):
    import torch
    rows = 100_000_000
    key, v1, v2 = workloads.make_groupby_table(torch, dev, rows)
    cap = capi.group_by_bound([capi.column(key)], rows)
    keys, aggs, g = capi.group_by([key], [v1, v2], workloads.AGGS7, rows, capacity=cap)
    hk, h1, h2 = key.cpu().numpy(), v1.cpu().numpy(), v2.cpu().numpy()
    ref = oracle.group_by_i64(hk, h1, h2, threads=THREADS, raw=True)
    del hk, h1, h2
    assert g == len(ref["key"]) and g > 900_000
    np.testing.assert_array_equal(keys[0][:g].cpu().numpy(), ref["key"])
    np.testing.assert_array_equal(aggs[0][:g].cpu().numpy(), ref["count"])
    np.testing.assert_array_equal(_hugeint_pairs(aggs[1], g), ref["sum_v1"])
    np.testing.assert_array_equal(aggs[2][:g].cpu().numpy(), ref["min_v1"])
    np.testing.assert_array_equal(aggs[3][:g].cpu().numpy(), ref["max_v1"])
    np.testing.assert_allclose(aggs[4][:g].cpu().numpy(), ref["sum_v2"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(aggs[5][:g].cpu().numpy(), ref["avg_v2"], rtol=1e-12, atol=0)
    assert int(ref["count"].sum()) == rows


def _probe_forms(prof):
    return {e["name"] for e in prof if e["name"].startswith("partition_probe_")}


def _join(fk, fpay, build, payload, with_probe_sum):
    if with_probe_sum:
        aggs = [(capi.SUM, 0, 0, capi.HUGEINT), (capi.SUM, 1, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64)]
        avals = [fpay]
    else:
        aggs = [(capi.SUM, 1, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64)]
        avals = []
    capi.profile_reset()
    capi.profile_enable(True)
    try:
        res = capi.join_aggregate(fk, avals, build, [payload], aggs)
    finally:
        capi.profile_enable(False)
    return res, _probe_forms(capi.profile_read())


def test_join_2p24_build_2p28_probe_closed_form(dev):
    import torch
    nb, nprobe = 1 << 24, 1 << 28
    build, payload, fk = workloads.make_join_tables(torch, dev, nb, nprobe)
    fpay = torch.empty(nprobe, dtype=torch.int32, device=dev)
    capi.generate_column(fpay, capi.INT32, 0, lo=-1000, hi=1000, seed=21, stream_id=5)
    hb, hp, hf = build.cpu().numpy(), payload.cpu().numpy().astype(np.int64), fk.cpu().numpy()
    assert np.array_equal(np.sort(hb), np.arange(nb))          # the build side is a permutation
    hit = hf < nb
    assert ((hf >= 0) & (hf < 2 * nb)).all()
    per_key = np.bincount(hf[hit], minlength=nb).astype(np.int64)  # probe rows per build key
    pay_by_key = np.empty(nb, np.int64)
    pay_by_key[hb] = hp                                        # payload of the build row holding key k
    count = int(hit.sum())
    total = int(np.dot(per_key, pay_by_key))
    probe_sum = int(fpay.cpu().numpy()[hit].astype(np.int64).sum())
    del hf

    # every table entry fits 16 bits: the LDS slice form (a full 65,536-entry slice per bucket)
    res, forms = _join(fk, fpay, build, payload, False)
    assert res == [total, count]
    assert forms == {"partition_probe_lds"}
    res, forms = _join(fk, fpay, build, payload, True)
    assert res == [probe_sum, total, count]
    assert forms == {"partition_probe_lds"}

    # one build payload past 12 bits: the 64-bit table through the caches, same closed form
    j = int(np.argmax(per_key[hb]))                            # a build row whose key is probed
    payload[j] = 5000
    total2 = total + int(per_key[hb[j]]) * (5000 - int(hp[j]))
    res, forms = _join(fk, fpay, build, payload, False)
    assert res == [total2, count]
    assert forms == {"partition_probe_l2"}
    res, forms = _join(fk, fpay, build, payload, True)
    assert res == [probe_sum, total2, count]
    assert forms == {"partition_probe_l2"}


def test_cross_4096_x_65536_filter(dev):
    import torch
    x, y = workloads.make_cross_tables(torch, dev)
    preds, projs = workloads.cross_exprs()
    (ox, oy), kept = capi.cross_filter([x], [y], preds, projs)
    hx, hy = x.cpu().numpy(), y.cpu().numpy()
    n, sx, sy = oracle.cross_filter_count(hx, hy, workloads.CROSS_LIMIT)
    assert kept == n and 0.45 * 4096 * 65536 < n < 0.55 * 4096 * 65536
    gx, gy = ox.cpu().numpy(), oy.cpu().numpy()
    assert int(gx.astype(np.int64).sum()) == sx and int(gy.astype(np.int64).sum()) == sy
    # left-major: left row i contributes, in right-row order, every y with x_i + y < 2^30
    per_left = np.searchsorted(np.sort(hy), workloads.CROSS_LIMIT - hx.astype(np.int64), side="left")
    assert int(per_left.sum()) == n
    starts = np.concatenate([[0], np.cumsum(per_left)])
    limit = workloads.CROSS_LIMIT
    for i in range(len(hx)):
        lo, hi = starts[i], starts[i + 1]
        assert (gx[lo:hi] == hx[i]).all()
        np.testing.assert_array_equal(gy[lo:hi], hy[hy.astype(np.int64) + int(hx[i]) < limit])


def test_batched_512_queries_against_oracle(dev):
    import torch
    cols, offsets, sizes, cards = workloads.make_batched_table(torch, dev, 512)
    seg, keys, aggs, goffs = capi.batched_filter_group_by(cols, offsets, workloads.BATCHED_PREDS, 2, [3, 4],
                                                          workloads.AGGS7)
    g = len(keys)
    assert len(goffs) == 513 and goffs[-1] == g
    hseg, hkeys = seg.cpu().numpy(), keys.cpu().numpy()
    count = aggs[0].cpu().numpy()
    s1 = _hugeint_pairs(aggs[1], g)
    mn, mx = aggs[2].cpu().numpy(), aggs[3].cpu().numpy()
    s2, avg = aggs[4].cpu().numpy(), aggs[5].cpu().numpy()
    a, c, key, v1, v2 = (t.cpu().numpy() for t in cols)
    offs = offsets.cpu().numpy()
    for q in range(512):
        r0, r1 = offs[q], offs[q + 1]
        keep = (a[r0:r1] < 500000) & (c[r0:r1] > 0.25)
        lo, hi = goffs[q], goffs[q + 1]
        assert (hseg[lo:hi] == q).all()
        ref = oracle.group_by_i64(key[r0:r1][keep], v1[r0:r1][keep], v2[r0:r1][keep], threads=THREADS, raw=True)
        assert hi - lo == len(ref["key"]), q
        np.testing.assert_array_equal(hkeys[lo:hi], ref["key"])
        np.testing.assert_array_equal(count[lo:hi], ref["count"])
        np.testing.assert_array_equal(s1[lo:hi], ref["sum_v1"])
        np.testing.assert_array_equal(mn[lo:hi], ref["min_v1"])
        np.testing.assert_array_equal(mx[lo:hi], ref["max_v1"])
        np.testing.assert_allclose(s2[lo:hi], ref["sum_v2"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(avg[lo:hi], ref["avg_v2"], rtol=1e-12, atol=0)
