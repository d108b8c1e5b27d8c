"""Sort-merge equi-join parity (aggregate-only and materialized) through the C ABI, against the oracle's
join and the closed form for permutation build keys. Integer results bit-exact; double sums within 1e-12."""
import numpy as np
import pytest

from vectisdb_amd import capi

pytestmark = pytest.mark.gpu


def _dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def config2_data(nb, np_, seed=0):
    rng = np.random.default_rng(seed)
    build = rng.permutation(nb).astype(np.int64)
    payload = rng.integers(0, 1000, nb).astype(np.int32)
    hit = rng.random(np_) < 0.9
    fk = np.where(hit, rng.integers(0, nb, np_), rng.integers(nb, 2 * nb, np_)).astype(np.int64)
    return build, payload, fk


@pytest.mark.parametrize("nb,np_", [(1, 1), (100, 1000), (1 << 16, 1 << 20), (1 << 20, 1 << 23)])
def test_join_sum_count_config2_shape(dev, nb, np_):
    import oracle
    build, payload, fk = config2_data(nb, np_, seed=nb)
    # A = probe (fk), B = build (key, payload): SELECT SUM(payload), COUNT(*)
    res = capi.join_aggregate(_dev(fk, dev), [], _dev(build, dev), [_dev(payload, dev)],
                              [(capi.SUM, 1, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64)])
    s, c = oracle.join_sum_count(build, payload, fk)
    assert res == [s, c]
    # closed form: build keys are a permutation of [0, nb)
    inv = np.empty(nb, np.int64)
    inv[build] = np.arange(nb)
    m = fk < nb
    assert c == int(m.sum())
    assert s == int(payload[inv[fk[m]]].astype(np.int64).sum())


def test_join_aggregate_duplicates_both_sides_and_probe_side_sums(dev):
    rng = np.random.default_rng(1)
    a = rng.integers(-50, 50, 20_000).astype(np.int64)
    b = rng.integers(-40, 60, 15_000).astype(np.int64)
    av = rng.integers(-10**6, 10**6, len(a)).astype(np.int64)
    bv = rng.standard_normal(len(b))
    res = capi.join_aggregate(_dev(a, dev), [_dev(av, dev)], _dev(b, dev), [_dev(bv, dev)],
                              [(capi.COUNT_STAR, -1, -1, capi.INT64), (capi.SUM, 0, 0, capi.HUGEINT),
                               (capi.SUM, 1, 0, capi.FLOAT64), (capi.MIN, 0, 0, capi.INT64),
                               (capi.MAX, 1, 0, capi.FLOAT64), (capi.AVG, 0, 0, capi.FLOAT64)])
    # brute force through per-key counts
    ka, ca = np.unique(a, return_counts=True)
    kb, cb = np.unique(b, return_counts=True)
    common = np.intersect1d(ka, kb)
    cnt = 0
    s_av = 0
    s_bv = []
    mn = None
    mx = None
    for k in common.tolist():
        na_ = int((a == k).sum())
        nb_ = int((b == k).sum())
        cnt += na_ * nb_
        s_av += int(av[a == k].sum()) * nb_
        s_bv += list(bv[b == k]) * na_
        mn = min(mn, int(av[a == k].min())) if mn is not None else int(av[a == k].min())
        mx = max(mx, float(bv[b == k].max())) if mx is not None else float(bv[b == k].max())
    import math
    exact_bv = math.fsum(s_bv)
    assert res[0] == cnt
    assert res[1] == s_av
    assert abs(res[2] - exact_bv) <= 1e-12 * abs(exact_bv)
    assert res[3] == mn and res[4] == mx
    assert abs(res[5] - s_av / cnt) <= 1e-12 * abs(s_av / cnt)


def test_join_aggregate_empty_and_disjoint(dev):
    a = np.arange(10, dtype=np.int64)
    b = np.arange(100, 110, dtype=np.int64)
    res = capi.join_aggregate(_dev(a, dev), [], _dev(b, dev), [], [(capi.COUNT_STAR, -1, -1, capi.INT64)])
    assert res == [0]


@pytest.mark.parametrize("seed,na,nb,hi", [(0, 1, 1, 1), (1, 1000, 800, 50), (2, 50_000, 40_000, 10_000),
                                           (3, 200_000, 3, 4), (4, 1_000_000, 800_000, 500_000)])
def test_join_pairs_match_oracle(dev, seed, na, nb, hi):
    import oracle
    rng = np.random.default_rng(seed)
    a = rng.integers(-hi, hi, na).astype(np.int64)
    b = rng.integers(-hi, hi, nb).astype(np.int64)
    ga, gb = capi.join_pairs(_dev(a, dev), _dev(b, dev))
    el, er = oracle.join_pairs(a, b)
    got = sorted(zip(ga.cpu().numpy().tolist(), gb.cpu().numpy().tolist()))
    assert got == sorted(zip(el.tolist(), er.tolist()))
    # output is ordered by key, then A row, then B row
    ka = a[ga.cpu().numpy()]
    assert np.all(np.diff(ka) >= 0)


def test_join_pairs_skewed_key(dev):
    """One key with 3,000 x 2,000 = 6M pairs is expanded across the device, not by one thread."""
    a = np.concatenate([np.full(3000, 7, np.int64), np.arange(100, 200, dtype=np.int64)])
    b = np.concatenate([np.full(2000, 7, np.int64), np.arange(150, 160, dtype=np.int64)])
    ga, gb = capi.join_pairs(_dev(a, dev), _dev(b, dev))
    assert len(ga) == 3000 * 2000 + 10
    assert np.all(a[ga.cpu().numpy()] == b[gb.cpu().numpy()])


def test_join_key_type_rules(dev):
    a = np.arange(10, dtype=np.int64)
    b = np.arange(10, dtype=np.int32)
    f = np.arange(10, dtype=np.float64)
    with pytest.raises(capi.VgpuError, match="same type"):
        capi.join_pairs(_dev(a, dev), _dev(b, dev))
    with pytest.raises(capi.VgpuError, match="integer/boolean"):
        capi.join_pairs(_dev(f, dev), _dev(f, dev))


@pytest.mark.parametrize("dtype,flip", [(np.int64, False), (np.int32, False), (np.int16, True), (np.int32, True)])
def test_join_count_side_narrowed_to_partner_range(dev, dtype, flip):
    """The larger side that only contributes counts is narrowed to the other side's key range before it is
    sorted (keys below and above that range, negative keys): COUNT(*), SUM over the other side and MIN/MAX
    stay exact against brute force, with the count-only side as A or as B."""
    rng = np.random.default_rng(int(np.dtype(dtype).itemsize) + flip)
    big = rng.integers(-3000, 9000, 200_000).astype(dtype)      # count-only side, wider range
    small = rng.integers(-1000, 5000, 20_000).astype(dtype)     # side with values
    vals = rng.integers(-10 ** 9, 10 ** 9, len(small)).astype(np.int64)
    kb, cb = np.unique(big, return_counts=True)
    ks, cs = np.unique(small, return_counts=True)
    count = 0
    total = 0
    mn = None
    for k, c in zip(kb.tolist(), cb.tolist()):
        sel = small == k
        if sel.any():
            count += c * int(sel.sum())
            total += c * int(vals[sel].sum())
            m = int(vals[sel].min())
            mn = m if mn is None else min(mn, m)
    if not flip:  # A = big (count only), B = small with values
        res = capi.join_aggregate(_dev(big, dev), [], _dev(small, dev), [_dev(vals, dev)],
                                  [(capi.COUNT_STAR, -1, -1, capi.INT64), (capi.SUM, 1, 0, capi.HUGEINT),
                                   (capi.MIN, 1, 0, capi.INT64)])
    else:         # A = small with values, B = big (count only)
        res = capi.join_aggregate(_dev(small, dev), [_dev(vals, dev)], _dev(big, dev), [],
                                  [(capi.COUNT_STAR, -1, -1, capi.INT64), (capi.SUM, 0, 0, capi.HUGEINT),
                                   (capi.MIN, 0, 0, capi.INT64)])
    assert res[0] == count and res[1] == total and res[2] == mn


def _brute_join(probe, build, payload):
    """COUNT(*) and SUM(payload) of probe JOIN build ON key, by per-key counts (exact Python ints)."""
    kp, cp = np.unique(probe, return_counts=True)
    count = 0
    total = 0
    for k, c in zip(kp.tolist(), cp.tolist()):
        sel = build == k
        if sel.any():
            count += c * int(sel.sum())
            if payload is not None:
                total += c * int(payload[sel].astype(np.int64).sum())
    return count, total


@pytest.mark.parametrize("dtype,flip,ptype", [(np.int64, False, np.int64), (np.int32, True, np.int64),
                                               (np.int16, False, np.int64), (np.int64, True, np.int32),
                                               (np.int32, False, np.int16), (np.int16, True, np.int32)])
def test_join_probe_form_duplicates_negative_keys_out_of_range(dev, dtype, flip, ptype):
    """Probe form (count-only side read once against a direct-address table of the other side): build keys
    with duplicates and negative values, 16/32/64-bit payloads, probe keys below, inside and above the build
    range; COUNT(*), SUM and AVG of the build payload exact against brute force, with the probe side as A or
    as B."""
    rng = np.random.default_rng(int(np.dtype(dtype).itemsize) * 2 + flip + 7 * np.dtype(ptype).itemsize)
    build = rng.integers(-2000, 6000, 30_000).astype(dtype)
    lim = 30_000 if ptype == np.int16 else 10 ** 6
    payload = rng.integers(-lim, lim, len(build)).astype(ptype)
    probe = rng.integers(-5000, 9000, 300_000).astype(dtype)
    count, total = _brute_join(probe, build, payload)
    if not flip:
        res = capi.join_aggregate(_dev(probe, dev), [], _dev(build, dev), [_dev(payload, dev)],
                                  [(capi.COUNT_STAR, -1, -1, capi.INT64), (capi.SUM, 1, 0, capi.HUGEINT),
                                   (capi.AVG, 1, 0, capi.FLOAT64)])
    else:
        res = capi.join_aggregate(_dev(build, dev), [_dev(payload, dev)], _dev(probe, dev), [],
                                  [(capi.COUNT_STAR, -1, -1, capi.INT64), (capi.SUM, 0, 0, capi.HUGEINT),
                                   (capi.AVG, 0, 0, capi.FLOAT64)])
    assert res[0] == count and res[1] == total
    assert abs(res[2] - total / count) <= 1e-12 * abs(total / count)


@pytest.mark.parametrize("case", ["count_only", "sparse_keys", "sum_does_not_pack", "int32_sum_does_not_pack",
                                  "long_runs"])
def test_join_probe_form_and_its_fallbacks(dev, case):
    """COUNT-only join (count-only table); build keys too sparse for a direct-address table and per-key sums
    past the table's 41-bit field both fall back to the sort-merge form; keys with more than 64 build rows
    build the table through the GROUP BY operator. All exact against brute force."""
    rng = np.random.default_rng(5)
    build = rng.integers(0, 20_000, 10_000).astype(np.int64)
    payload = rng.integers(0, 1000, len(build)).astype(np.int64)
    if case == "long_runs":  # ~100 build rows per key
        build = rng.integers(-50, 150, 20_000).astype(np.int64)
        payload = rng.integers(-1000, 1000, len(build)).astype(np.int64)
    if case == "sparse_keys":
        build = build * 100_003 - 7   # span ~2e9, far above 2 x rows + 2^20
    if case == "sum_does_not_pack":
        payload = payload + (1 << 45)
    if case == "int32_sum_does_not_pack":  # 10K INT32 rows near 2^31 over 8 keys: per-key sums past 2^40
        build = build % 8
        payload = (payload + (1 << 31) - 1000).astype(np.int32)
    probe = rng.choice(np.concatenate([build, build + 1]), 200_000).astype(np.int64)
    count, total = _brute_join(probe, build, None if case == "count_only" else payload)
    if case == "count_only":
        res = capi.join_aggregate(_dev(probe, dev), [], _dev(build, dev), [],
                                  [(capi.COUNT_STAR, -1, -1, capi.INT64)])
        assert res == [count]
    else:
        res = capi.join_aggregate(_dev(probe, dev), [], _dev(build, dev), [_dev(payload, dev)],
                                  [(capi.COUNT_STAR, -1, -1, capi.INT64), (capi.SUM, 1, 0, capi.HUGEINT)])
        assert res == [count, total]


@pytest.mark.parametrize("dtype,skew,flip,wide", [(np.int64, False, False, False), (np.int32, True, False, False),
                                                  (np.int64, True, True, False), (np.int64, False, False, True),
                                                  (np.int32, True, True, True)])
def test_join_partitioned_probe_form(dev, dtype, skew, flip, wide):
    """A probe side at least as large as a direct-address table above 4 MB goes through the key-range
    partition (histogram, scatter into 256 buckets, bucket-ordered probe): from 16-bit table slices in LDS
    when every entry fits (payloads below 2^11), from the 64-bit table in L2 otherwise (wide payloads).
    Build keys: a shifted permutation (negative keys included), so COUNT(*) and SUM(payload) have a closed
    form; probe keys below, inside and above the table, uniform or with one key repeated 2^21 times (one
    bucket far larger than the rest)."""
    nb = 1 << 20
    shift = 300_000
    rng = np.random.default_rng(41 + skew + 2 * flip + 4 * wide)
    build = (rng.permutation(nb) - shift).astype(dtype)
    payload = rng.integers(-10 ** 6 if wide else 0, 10 ** 6 if wide else 1000, nb).astype(np.int32)
    hits = rng.integers(0, nb, 3 << 20) - shift
    below = rng.integers(-2 * nb, -shift, 1 << 18)
    above = rng.integers(nb - shift, 3 * nb, 1 << 18)
    parts = [hits, below, above]
    if skew:
        parts.append(np.full(1 << 21, int(build[7]), dtype=np.int64))
    probe = rng.permutation(np.concatenate(parts)).astype(dtype)
    inv = np.empty(nb, np.int64)
    inv[build.astype(np.int64) + shift] = np.arange(nb)
    idx = probe.astype(np.int64) + shift
    m = (idx >= 0) & (idx < nb)
    count = int(m.sum())
    total = int(payload[inv[idx[m]]].astype(np.int64).sum())
    if not flip:
        res = capi.join_aggregate(_dev(probe, dev), [], _dev(build, dev), [_dev(payload, dev)],
                                  [(capi.SUM, 1, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64)])
    else:
        res = capi.join_aggregate(_dev(build, dev), [_dev(payload, dev)], _dev(probe, dev), [],
                                  [(capi.SUM, 0, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64)])
    assert res == [total, count]


def _brute_probe_side_sums(probe, py, build, bpay):
    """COUNT(*), SUM(build payload), SUM(probe payload) of probe JOIN build (exact Python ints)."""
    ub, cb = np.unique(build, return_counts=True)
    sums_b = {}
    if bpay is not None:
        order = np.argsort(build, kind="stable")
        sb = np.add.reduceat(bpay[order].astype(object), np.searchsorted(build[order], ub)) if len(ub) else []
        sums_b = dict(zip(ub.tolist(), [int(x) for x in sb]))
    pos = np.searchsorted(ub, probe)
    pos = np.minimum(pos, len(ub) - 1)
    hit = ub[pos] == probe
    c = np.where(hit, cb[pos], 0).astype(np.int64)
    count = int(c.sum())
    sum_y = int(np.dot(py.astype(object), c.astype(object))) if py is not None else 0
    sum_b = sum(sums_b.get(k, 0) for k in probe[hit].tolist()) if bpay is not None else 0
    return count, sum_b, sum_y


@pytest.mark.parametrize("ktype,ptype,wide_build,flip", [(np.int64, np.int64, False, False),
                                                         (np.int32, np.int32, False, True),
                                                         (np.int64, np.int16, True, False),
                                                         (np.int16, np.int64, False, True)])
def test_join_probe_form_larger_side_sums(dev, ktype, ptype, wide_build, flip):
    """SUM/AVG of a column of the larger (probe) side, with and without a SUM of the build side's payload:
    the probe payload travels with its key through the key-range partition and each probe row adds its
    payload times its key's build row count (64 x 64 -> 128-bit products). Duplicate and negative build
    keys; both table forms (16-bit slices in LDS, 64-bit in L2 for wide build payloads); the probe side as
    A or as B. Exact against brute force."""
    rng = np.random.default_rng(int(np.dtype(ktype).itemsize) * 10 + int(np.dtype(ptype).itemsize) + 3 * flip)
    lo, hi = (-3000, 9000) if ktype == np.int16 else (-200_000, 600_000)
    build = rng.integers(lo // 2, hi // 2, 150_000).astype(ktype)
    bpay = rng.integers(-10 ** 6 if wide_build else 0, 10 ** 6 if wide_build else 500, len(build)).astype(np.int32)
    probe = rng.integers(lo, hi, 1_200_000).astype(ktype)
    ylim = {np.int16: 30_000, np.int32: 2 ** 31 - 1, np.int64: 2 ** 62}[ptype]
    py = rng.integers(-ylim, ylim, len(probe)).astype(ptype)
    count, sum_b, sum_y = _brute_probe_side_sums(probe.astype(np.int64), py, build.astype(np.int64), bpay)
    if not flip:  # A = probe (larger) with y, B = build with payload
        res = capi.join_aggregate(_dev(probe, dev), [_dev(py, dev)], _dev(build, dev), [_dev(bpay, dev)],
                                  [(capi.SUM, 0, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64),
                                   (capi.SUM, 1, 0, capi.HUGEINT), (capi.AVG, 0, 0, capi.FLOAT64)])
        only_y = capi.join_aggregate(_dev(probe, dev), [_dev(py, dev)], _dev(build, dev), [],
                                     [(capi.SUM, 0, 0, capi.HUGEINT)])
    else:         # A = build with payload, B = probe (larger) with y
        res = capi.join_aggregate(_dev(build, dev), [_dev(bpay, dev)], _dev(probe, dev), [_dev(py, dev)],
                                  [(capi.SUM, 1, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64),
                                   (capi.SUM, 0, 0, capi.HUGEINT), (capi.AVG, 1, 0, capi.FLOAT64)])
        only_y = capi.join_aggregate(_dev(build, dev), [], _dev(probe, dev), [_dev(py, dev)],
                                     [(capi.SUM, 1, 0, capi.HUGEINT)])
    assert res[0] == sum_y and res[1] == count and res[2] == sum_b
    assert abs(res[3] - sum_y / count) <= 1e-12 * abs(sum_y / count)
    assert only_y == [sum_y]
