"""GROUP BY parity: the HIP radix-sort + segmented-reduce path (through the C ABI) against the oracle.

Bar: bit-exact for keys, COUNT, SUM(int) (HUGEINT), MIN, MAX; DOUBLE SUM/AVG within 1e-12 relative of the
correctly rounded sum (the floating-point tolerance of this path; the engine's double-double accumulation
is normally exact).
"""
import numpy as np
import pytest

from vectisdb_amd import capi

pytestmark = pytest.mark.gpu

AGGS7 = [
    (capi.COUNT_STAR, -1, capi.INT64),
    (capi.SUM, 0, capi.HUGEINT),
    (capi.MIN, 0, capi.INT64),
    (capi.MAX, 0, capi.INT64),
    (capi.SUM, 1, capi.FLOAT64),
    (capi.AVG, 1, capi.FLOAT64),
]


def _hugeint_to_ints(t):
    a = t.cpu().numpy().reshape(-1, 2)
    return [int(lo) % (1 << 64) + (int(hi) << 64) for lo, hi in a]


def _run(dev, key, v1, v2):
    import torch
    k = torch.from_numpy(key).to(dev)
    a = torch.from_numpy(v1).to(dev)
    b = torch.from_numpy(v2).to(dev)
    keys, aggs, g = capi.group_by([k], [a, b], AGGS7, len(key))
    return keys, aggs, g


def _check(dev, key, v1, v2):
    keys, aggs, g = _run(dev, key, v1, v2)
    ref = capi_ref = None
    import oracle
    ref = oracle.group_by_i64(key, v1, v2)
    assert g == len(ref["key"])
    gk = keys[0].cpu().numpy()
    order = np.argsort(gk, kind="stable")  # groups come out in ascending key order; sort defensively
    np.testing.assert_array_equal(gk[order], ref["key"])
    np.testing.assert_array_equal(aggs[0].cpu().numpy()[order], ref["count"])
    s1 = _hugeint_to_ints(aggs[1])
    assert [s1[i] for i in order] == ref["sum_v1"]
    np.testing.assert_array_equal(aggs[2].cpu().numpy()[order], ref["min_v1"])
    np.testing.assert_array_equal(aggs[3].cpu().numpy()[order], ref["max_v1"])
    s2 = aggs[4].cpu().numpy()[order]
    avg = aggs[5].cpu().numpy()[order]
    tol = 1e-12
    np.testing.assert_allclose(s2, ref["sum_v2"], rtol=tol, atol=0)
    np.testing.assert_allclose(avg, ref["avg_v2"], rtol=tol, atol=0)
    return g


@pytest.mark.parametrize("n,distinct", [(1, 1), (7, 3), (2048, 17), (2049, 2049), (100_000, 1000),
                                        (1_000_003, 50_000), (3_000_000, 1_000_000)])
def test_groupby_uniform_keys(dev, n, distinct):
    rng = np.random.default_rng(n)
    key = rng.integers(0, distinct, n, dtype=np.int64)
    v1 = rng.integers(-1_000_000, 1_000_001, n, dtype=np.int64)
    v2 = rng.standard_normal(n)
    _check(dev, key, v1, v2)


def test_groupby_zipf_skew(dev):
    """A few keys own most rows: groups span many tiles, exercising the tail/lead fix-up."""
    rng = np.random.default_rng(7)
    n = 4_000_000
    ranks = rng.zipf(1.1, n)
    ranks = ranks[ranks <= 1_000_000]
    perm = rng.permutation(1_000_000).astype(np.int64)
    key = perm[ranks - 1]
    n = len(key)
    v1 = rng.integers(-1_000_000, 1_000_001, n, dtype=np.int64)
    v2 = rng.standard_normal(n)
    _check(dev, key, v1, v2)


def test_groupby_single_key_all_rows(dev):
    n = 1_500_000
    key = np.full(n, 42, np.int64)
    rng = np.random.default_rng(3)
    _check(dev, key, rng.integers(-5, 5, n, dtype=np.int64), rng.standard_normal(n))


def test_groupby_negative_and_wide_keys(dev):
    rng = np.random.default_rng(11)
    n = 500_000
    key = rng.choice(np.array([-2**62, -1, 0, 1, 2**40, 2**62 - 1], np.int64), n)
    key[::7] = rng.integers(-2**63, 2**63 - 1, len(key[::7]), dtype=np.int64)
    _check(dev, key, rng.integers(-10**6, 10**6, n, dtype=np.int64), rng.standard_normal(n))


def test_groupby_sorted_and_reverse_input(dev):
    n = 300_000
    key = np.repeat(np.arange(1000, dtype=np.int64), n // 1000)
    rng = np.random.default_rng(5)
    v1 = rng.integers(-10**6, 10**6, len(key), dtype=np.int64)
    v2 = rng.standard_normal(len(key))
    _check(dev, key, v1, v2)
    _check(dev, key[::-1].copy(), v1, v2)


def test_groupby_global_aggregate(dev):
    import torch
    rng = np.random.default_rng(9)
    n = 777_777
    v1 = rng.integers(-10**6, 10**6, n, dtype=np.int64)
    v2 = rng.standard_normal(n)
    a = torch.from_numpy(v1).to(dev)
    b = torch.from_numpy(v2).to(dev)
    _, aggs, g = capi.group_by([], [a, b], AGGS7, n)
    assert g == 1
    assert int(aggs[0].cpu()[0]) == n
    assert _hugeint_to_ints(aggs[1])[0] == int(v1.sum())
    assert int(aggs[2].cpu()[0]) == v1.min() and int(aggs[3].cpu()[0]) == v1.max()
    import math
    exact = math.fsum(v2)
    assert abs(float(aggs[4].cpu()[0]) - exact) <= 1e-12 * abs(exact)


def test_groupby_int32_keys_and_narrow_minmax(dev):
    import torch
    rng = np.random.default_rng(13)
    n = 200_000
    key = rng.integers(-50, 50, n).astype(np.int32)
    v = rng.integers(-30000, 30000, n).astype(np.int16)
    keys, aggs, g = capi.group_by([torch.from_numpy(key).to(dev)], [torch.from_numpy(v).to(dev)],
                                  [(capi.MIN, 0, capi.INT16), (capi.MAX, 0, capi.INT16), (capi.COUNT, 0, capi.INT64),
                                   (capi.SUM, 0, capi.HUGEINT)], n)
    gk = keys[0].cpu().numpy()
    uk = np.unique(key)
    np.testing.assert_array_equal(np.sort(gk), uk)
    mins = aggs[0].cpu().numpy()
    maxs = aggs[1].cpu().numpy()
    cnts = aggs[2].cpu().numpy()
    sums = _hugeint_to_ints(aggs[3])
    for i, k in enumerate(gk):
        sel = v[key == k]
        assert mins[i] == sel.min() and maxs[i] == sel.max() and cnts[i] == len(sel)
        assert sums[i] == int(sel.astype(np.int64).sum())


def test_groupby_two_key_columns(dev):
    import torch
    rng = np.random.default_rng(17)
    n = 400_000
    k1 = rng.integers(0, 40, n).astype(np.int32)
    k2 = rng.integers(-1000, 1000, n).astype(np.int64)
    v = rng.integers(-10**6, 10**6, n).astype(np.int64)
    keys, aggs, g = capi.group_by([torch.from_numpy(k1).to(dev), torch.from_numpy(k2).to(dev)],
                                  [torch.from_numpy(v).to(dev)],
                                  [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT)], n)
    got = {}
    c = aggs[0].cpu().numpy()
    s = _hugeint_to_ints(aggs[1])
    for a, b, cc, ss in zip(keys[0].cpu().numpy(), keys[1].cpu().numpy(), c, s):
        got[(int(a), int(b))] = (int(cc), ss)
    exp = {}
    for a, b, vv in zip(k1.tolist(), k2.tolist(), v.tolist()):
        cc, ss = exp.get((a, b), (0, 0))
        exp[(a, b)] = (cc + 1, ss + vv)
    assert got == exp
    # groups come out in lexicographic (k1, k2) order
    pairs = list(zip(keys[0].cpu().numpy().tolist(), keys[1].cpu().numpy().tolist()))
    assert pairs == sorted(pairs)


def test_groupby_empty_input(dev):
    import torch
    k = torch.empty(0, dtype=torch.int64, device=dev)
    v = torch.empty(0, dtype=torch.int64, device=dev)
    keys, aggs, g = capi.group_by([k], [v], [(capi.COUNT_STAR, -1, capi.INT64)], 0, capacity=1)
    assert g == 0


def test_groupby_rejects_float_key_and_bad_types(dev):
    import torch
    f = torch.zeros(10, dtype=torch.float64, device=dev)
    i = torch.zeros(10, dtype=torch.int64, device=dev)
    with pytest.raises(capi.VgpuError, match="integer/boolean"):
        capi.group_by([f], [i], [(capi.COUNT_STAR, -1, capi.INT64)], 10, capacity=10)
    with pytest.raises(capi.VgpuError, match="MIN/MAX output type"):
        capi.group_by([i], [i], [(capi.MIN, 0, capi.INT32)], 10, capacity=10)
    with pytest.raises(capi.VgpuError, match="AVG must produce DOUBLE"):
        capi.group_by([i], [i], [(capi.AVG, 0, capi.INT64)], 10, capacity=10)


def test_sort_column_matches_stable_argsort(dev):
    import torch
    rng = np.random.default_rng(21)
    for n, hi in [(1, 5), (1000, 3), (65_537, 1 << 20), (2_000_000, 1 << 40)]:
        key = rng.integers(-hi, hi, n, dtype=np.int64)
        out, perm = capi.sort_column(torch.from_numpy(key).to(dev))
        order = np.argsort(key, kind="stable")
        np.testing.assert_array_equal(out.cpu().numpy(), key[order])
        np.testing.assert_array_equal(perm.cpu().numpy().astype(np.int64), order)


def test_filter_gt_i32(dev):
    """FilterGreaterThanInt32: the kept values in input order, for sizes around the 8,192-row tile and
    extreme thresholds (INT32_MAX keeps nothing, INT32_MIN everything but INT32_MIN)."""
    import torch
    rng = np.random.default_rng(23)
    lo32, hi32 = np.iinfo(np.int32).min, np.iinfo(np.int32).max
    for n in [0, 1, 2047, 2048, 8191, 8192, 8193, 1_000_001, (1 << 24) + 5]:
        x = rng.integers(-100, 100, n).astype(np.int32)
        got = capi.filter_gt_i32(torch.from_numpy(x).to(dev), 10)
        np.testing.assert_array_equal(got.cpu().numpy(), x[x > 10])
    x = rng.integers(lo32, hi32, 100_003, dtype=np.int64, endpoint=True).astype(np.int32)
    x[:3] = [lo32, hi32, 0]
    xd = torch.from_numpy(x).to(dev)
    for t in [hi32, lo32, -1, 0, hi32 - 1]:
        got = capi.filter_gt_i32(xd, int(t))
        np.testing.assert_array_equal(got.cpu().numpy(), x[x > t])


def test_groupby_under_sort_autotuner(dev):
    """With the autotuner on, the first sort of a shape runs every radix-pass variant and keeps the fastest;
    the results stay bit-exact against the oracle on the tuning call and on the cached-choice call, and the
    shape appears in the report with a valid choice."""
    rng = np.random.default_rng(23)
    n = 1_500_000
    key = rng.integers(0, 300_000, n).astype(np.int64)
    v1 = rng.integers(-10 ** 6, 10 ** 6 + 1, n).astype(np.int64)
    v2 = rng.standard_normal(n)
    capi.sort_autotune(True)
    try:
        g1 = _check(dev, key, v1, v2)   # tunes
        g2 = _check(dev, key, v1, v2)   # uses the cached choice
        assert g1 == g2
        report = capi.sort_autotune_report()
        mine = [r for r in report if r["npayload"] == 2 and r["log2rows"] == 20]
        assert mine, report
        for r in mine:
            assert r["radix_bits"] in (8, 11) and r["items"] in (16, 32) and r["prefetch"] in (0, 1)
            assert r["block"] in (256, 512, 1024)
            assert r["candidates"] >= 6 and r["ms"] > 0
    finally:
        capi.sort_autotune(False)


def test_prepared_group_by_matches_group_by(dev):
    """The prepared call (argument arrays built once, as the benchmark's timed loop uses it) returns the same
    groups as the plain call, on every one of several consecutive calls."""
    import torch
    rng = np.random.default_rng(31)
    n = 300_000
    key = rng.integers(-5_000, 5_000, n).astype(np.int64)
    v1 = rng.integers(-10 ** 6, 10 ** 6 + 1, n).astype(np.int64)
    v2 = rng.standard_normal(n)
    kd, ad, bd = (torch.from_numpy(x).to(dev) for x in (key, v1, v2))
    keys, aggs, g = capi.group_by([kd], [ad, bd], AGGS7, n)
    cap = capi.group_by_bound([capi.column(kd)], n)
    outs = capi.group_by_outputs([kd], AGGS7, cap)
    step = capi.PreparedGroupBy([kd], [ad, bd], AGGS7, n, cap, outs)
    for _ in range(3):
        assert step() == g
        pk, pa, pg = step.result()
        assert pg == g
        assert torch.equal(pk[0], keys[0])
        for a, b in zip(pa, aggs):
            assert torch.equal(a, b)


def test_profile_only_times_the_named_kernel(dev):
    """With a profile name filter, only launches under that name are bracketed and reported (the benchmark
    times the dominant kernel this way inside its timed loop); without it every launch is."""
    import torch
    rng = np.random.default_rng(37)
    n = 1 << 20
    kd = torch.from_numpy(rng.integers(0, 1 << 18, n).astype(np.int64)).to(dev)
    vd = torch.from_numpy(rng.integers(-100, 100, n).astype(np.int64)).to(dev)
    aggs = [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT)]
    try:
        capi.profile_reset()
        capi.profile_enable(True)
        capi.group_by([kd], [vd], aggs, n)
        every = {e["name"] for e in capi.profile_read()}
        assert "radix_pass_first" in every and len(every) > 1
        capi.profile_reset()
        capi.profile_only("radix_pass_first")
        capi.group_by([kd], [vd], aggs, n)
        capi.group_by([kd], [vd], aggs, n)
        only = capi.profile_read()
        assert [e["name"] for e in only] == ["radix_pass_first"] and only[0]["launches"] == 2
    finally:
        capi.profile_enable(False)
        capi.profile_only(None)


def test_autotuned_shape_reused_with_other_call_settings(dev):
    """The autotuner caches a pass variant per sort shape, not the settings of the call that tuned it: a later
    GROUP BY of the same shape with an aggregate mix the fused last pass does not cover must take the plain
    last pass, and one whose INT64 column holds a value outside INT32 must flag its own narrowing overflow and
    sort again with the column whole. Each result against numpy."""
    import torch
    rng = np.random.default_rng(29)
    n = 2_000_000
    # 17-bit keys: three 8-bit passes whose last digit has 2 values, so at most ~132K partial states (15x fewer
    # than rows) and the fused last pass applies to the seven-aggregate mix
    key = rng.integers(0, 1 << 17, n).astype(np.int64)
    v1 = rng.integers(-10 ** 6, 10 ** 6 + 1, n).astype(np.int64)
    v2 = rng.standard_normal(n)
    order = np.argsort(key, kind="stable")
    uk, first = np.unique(key[order], return_index=True)
    capi.sort_autotune(True)
    try:
        capi.profile_reset()
        capi.profile_enable(True)
        _check(dev, key, v1, v2)  # tunes the shape; the fused last pass applies to this mix
        capi.profile_enable(False)
        assert any(e["name"] == "radix_pass_reduce" for e in capi.profile_read())
        kd, ad, bd = (torch.from_numpy(x).to(dev) for x in (key, v1, v2))
        capi.profile_reset()
        capi.profile_enable(True)
        keys, aggs, g = capi.group_by([kd], [ad, bd], [(capi.MIN, 1, capi.FLOAT64), (capi.MAX, 0, capi.INT64)], n)
        capi.profile_enable(False)
        assert not any(e["name"] == "radix_pass_reduce" for e in capi.profile_read())
        assert g == len(uk)
        np.testing.assert_array_equal(keys[0].cpu().numpy(), uk)
        np.testing.assert_array_equal(aggs[0].cpu().numpy(), np.minimum.reduceat(v2[order], first))
        np.testing.assert_array_equal(aggs[1].cpu().numpy(), np.maximum.reduceat(v1[order], first))
        wide = v1.copy()
        wide[n // 2] = 1 << 40
        _check(dev, key, wide, v2)
    finally:
        capi.sort_autotune(False)
        capi.profile_enable(False)


@pytest.mark.parametrize("variant", [
    {"VGPU_SORT_BLOCK": "512"},
    {"VGPU_SORT_BLOCK": "1024"},
    {"VGPU_SORT_ITEMS": "32", "VGPU_SORT_PREFETCH": "0"},
    {"VGPU_SORT_ITEMS": "32", "VGPU_SORT_PREFETCH": "1"},
    {"VGPU_SORT_ITEMS": "16", "VGPU_SORT_PREFETCH": "0"},
    {"VGPU_SORT_ITEMS": "16", "VGPU_SORT_PREFETCH": "1"},
], ids=["b512_i16", "b1024_i16", "i32_load_late", "i32_prefetch", "i16_load_late", "i16_prefetch"])
def test_sort_variants_forced(dev, monkeypatch, variant):
    """Every 8-bit radix-pass variant (threads per tile x items per thread x payload prefetch), forced through
    the tuning overrides: GROUP BY bit-exact against the oracle with row counts around the 4096- and
    8192-row tile edges (all at least 64K rows, where 8-bit digits are used), one-pass and three-pass keys,
    negative keys; and the key-only sort with its permutation against a stable argsort."""
    import torch
    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(31)
    for n, distinct in [(65_537, 600), (131_071, 131_071), (135_169, 5), (2_500_001, 1 << 20)]:
        key = rng.integers(-distinct, distinct, n, dtype=np.int64)
        v1 = rng.integers(-10 ** 9, 10 ** 9 + 1, n, dtype=np.int64)
        v2 = rng.standard_normal(n)
        _check(dev, key, v1, v2)
    for n in (65_537, 1_000_003):
        key = rng.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
        out, perm = capi.sort_column(torch.from_numpy(key).to(dev))
        order = np.argsort(key, kind="stable")
        np.testing.assert_array_equal(out.cpu().numpy(), key[order])
        np.testing.assert_array_equal(perm.cpu().numpy().astype(np.int64), order)


def test_groupby_many_tiles_carries(dev):
    """20M rows = 9,766 reduce tiles, so the tiles' group offsets come from more than one first-level scan
    block (4,096 tiles each) plus carries; Zipf keys put long groups across tile and scan-block edges."""
    rng = np.random.default_rng(29)
    n = 20_000_000
    ranks = rng.zipf(1.2, n)
    ranks = ranks[ranks <= 200_000]
    key = (rng.permutation(200_000).astype(np.int64) * 3 - 100_000)[ranks - 1]
    n = len(key)
    v1 = rng.integers(-1_000_000, 1_000_001, n, dtype=np.int64)
    v2 = rng.standard_normal(n)
    _check(dev, key, v1, v2)


def test_groupby_count_only_many_tiles(dev):
    """COUNT-only GROUP BY (the run-length path) over 20M rows: keys and counts equal numpy's unique."""
    import torch
    rng = np.random.default_rng(31)
    n = 20_000_000
    key = rng.integers(-3_000_000, 3_000_000, n, dtype=np.int64)
    key[: n // 4] = 17  # one group spanning thousands of tiles
    k = torch.from_numpy(key).to(dev)
    keys, aggs, g = capi.group_by([k], [], [(capi.COUNT_STAR, -1, capi.INT64)], n)
    uk, uc = np.unique(key, return_counts=True)
    assert g == len(uk)
    gk = keys[0].cpu().numpy()[:g]
    np.testing.assert_array_equal(gk, uk)
    np.testing.assert_array_equal(aggs[0].cpu().numpy()[:g], uc)


@pytest.mark.parametrize("n", [1, 5, 2047, 2048, 2049, 4101, 100_003])
@pytest.mark.parametrize("span_bits,distinct", [(30, 0), (30, 7), (40, 0), (40, 50)])
def test_groupby_count_only_run_length_edges(dev, n, span_bits, distinct):
    """COUNT-only GROUP BY through the run-length path around its 2,048-row tile and 256-row step edges, with
    32-bit (30-bit span) and 64-bit (40-bit span) packed keys, all distinct or a few heavy groups: keys and
    counts equal numpy's unique."""
    import torch
    rng = np.random.default_rng(n * 7 + span_bits + distinct)
    if distinct:
        vals = rng.integers(-(1 << (span_bits - 1)), 1 << (span_bits - 1), distinct, dtype=np.int64)
        key = vals[rng.integers(0, distinct, n)]
    else:
        key = rng.integers(-(1 << (span_bits - 1)), 1 << (span_bits - 1), n, dtype=np.int64)
    k = torch.from_numpy(key).to(dev)
    keys, aggs, g = capi.group_by([k], [], [(capi.COUNT_STAR, -1, capi.INT64)], n)
    uk, uc = np.unique(key, return_counts=True)
    assert g == len(uk)
    np.testing.assert_array_equal(keys[0].cpu().numpy()[:g], uk)
    np.testing.assert_array_equal(aggs[0].cpu().numpy()[:g], uc)


def test_groupby_int64_values_at_int32_bounds(dev):
    """INT64 aggregate inputs are carried through the sort as INT32 when they fit: values exactly at
    INT32_MIN / INT32_MAX fit, so SUM (HUGEINT), MIN and MAX come from the narrowed carry and must still be
    exact."""
    rng = np.random.default_rng(41)
    n = 3_000_000
    key = rng.integers(0, 20_000, n, dtype=np.int64)
    v1 = rng.integers(-(1 << 31), 1 << 31, n, dtype=np.int64)
    v1[::7] = -(1 << 31)
    v1[3::11] = (1 << 31) - 1
    v2 = rng.standard_normal(n)
    _check(dev, key, v1, v2)


@pytest.mark.parametrize("where", ["first", "middle", "last"])
def test_groupby_int64_values_beyond_int32(dev, where):
    """One value just outside INT32 anywhere in the column makes the narrowed sort report an overflow; the
    GROUP BY then runs again carrying the column whole. The result must equal the oracle's, and so must a
    second call over the same column (remembered as wide) and a call over a column that fits."""
    import torch
    import oracle
    rng = np.random.default_rng({"first": 43, "middle": 47, "last": 53}[where])
    n = 2_500_000
    key = rng.integers(-5_000, 5_000, n, dtype=np.int64)
    v1 = rng.integers(-1_000_000, 1_000_001, n, dtype=np.int64)
    row = {"first": 0, "middle": n // 2 + 123, "last": n - 1}[where]
    v1[row] = (1 << 31) if where != "middle" else -(1 << 31) - 1
    v1[rng.integers(0, n, 5)] = rng.integers(-(1 << 62), 1 << 62, 5)
    v2 = rng.standard_normal(n)
    _check(dev, key, v1, v2)
    _check(dev, key, v1, v2)
    fits = np.clip(v1, -(1 << 31), (1 << 31) - 1)
    _check(dev, key, fits, v2)
    # an executor-style call with COUNT(col)-free MIN/MAX only over the wide column
    k = torch.from_numpy(key).to(dev)
    a = torch.from_numpy(v1).to(dev)
    keys, aggs, g = capi.group_by([k], [a], [(capi.MIN, 0, capi.INT64), (capi.MAX, 0, capi.INT64)], n)
    ref = oracle.group_by_i64(key, v1, v2)
    assert g == len(ref["key"])
    np.testing.assert_array_equal(aggs[0].cpu().numpy(), ref["min_v1"])
    np.testing.assert_array_equal(aggs[1].cpu().numpy(), ref["max_v1"])
