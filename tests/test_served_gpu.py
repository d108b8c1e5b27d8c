"""Served form: a mixed burst of filter/projection, GROUP BY and join-aggregate queries over row ranges of
device-resident tables, run through a batching server (continuous batching, 4 executors) and through the
plain per-query server (one query at a time). Each query's result must be the same alone and in the mix --
bit-exact for rows, keys, COUNT, integer SUM/MIN/MAX and join SUM/COUNT, DOUBLE SUM/AVG within 1e-12 --
and both must match the oracle."""
import numpy as np
import pytest

import oracle
from vectisdb_amd.workloads import cfg0_exprs
from vectisdb_amd import capi

pytestmark = pytest.mark.gpu

GB_AGGS = [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT), (capi.MIN, 0, capi.INT64),
           (capi.MAX, 0, capi.INT64), (capi.SUM, 1, capi.FLOAT64), (capi.AVG, 1, capi.FLOAT64)]


CROSS_PRED = ("((bool*)out)[idx] = (((const int*)inputs[0])[row] + ((const int*)inputs[1])[row]) < 1073741824;",
              [0, 1], capi.BOOLEAN)
CROSS_PROJS = [("((int*)out)[idx] = ((const int*)inputs[0])[row];", [0], capi.INT32),
               ("((int*)out)[idx] = ((const int*)inputs[0])[row];", [1], capi.INT32)]


def hugeints(a):
    return [int(lo) % (1 << 64) + (int(hi) << 64) for lo, hi in a]


@pytest.fixture(scope="module")
def tables(dev):
    import torch
    rng = np.random.default_rng(5)
    n = 1 << 21
    a = rng.integers(0, 10 ** 6, n).astype(np.int32)
    b = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    c = rng.random(n)
    key = rng.integers(0, 5000, n).astype(np.int64)
    v1 = rng.integers(-10 ** 6, 10 ** 6 + 1, n).astype(np.int64)
    v2 = rng.standard_normal(n)
    nb = 200_000
    bkey = np.concatenate([rng.permutation(nb), rng.integers(0, nb // 10, 20_000)]).astype(np.int64)  # duplicates
    payload = rng.integers(0, 1000, len(bkey)).astype(np.int32)
    fk = np.where(rng.random(n) < 0.9, rng.integers(0, nb, n), rng.integers(nb, 2 * nb, n)).astype(np.int64)
    x = rng.integers(0, 1 << 30, 4096).astype(np.int32)
    y = rng.integers(0, 1 << 30, 65536).astype(np.int32)
    host = dict(a=a, b=b, c=c, key=key, v1=v1, v2=v2, bkey=bkey, payload=payload, fk=fk, x=x, y=y)
    devt = {k: torch.from_numpy(v).to(dev) for k, v in host.items()}
    torch.cuda.synchronize()
    return host, devt


def add_templates(server, d):
    preds, projs = cfg0_exprs()
    t_fp = server.add_filter_project([d["a"], d["b"], d["c"]], preds, projs)
    t_gb = server.add_group_by([d["key"], d["v1"], d["v2"]], [], 0, [1, 2], GB_AGGS)
    t_j = server.add_join_agg(d["fk"], d["bkey"], d["payload"])
    t_x = server.add_cross([d["x"]], [d["y"]], [CROSS_PRED], CROSS_PROJS)
    return t_fp, t_gb, t_j, t_x


def make_queries(n_rows, nq, seed, n_right=65536):
    rng = np.random.default_rng(seed)
    q = []
    for i in range(nq):
        kind = i % 4
        size = int(np.exp(rng.uniform(np.log(1), np.log(400_000))))
        if kind == 3:
            size = size % 2500  # right-side rows of a cross query (x 4,096 left rows)
        if i in (4, 5, 6, 7):
            size = 0  # empty ranges of every template
        if i in (8, 9, 10, 11):
            size = 1
        rows = n_right if kind == 3 else n_rows
        start = int(rng.integers(0, rows - size + 1))
        q.append((kind, start, size))
    q.append((1, 0, n_rows))  # a whole-table GROUP BY in the mix
    q.append((0, 100, 4093))  # unaligned start, ragged length
    return q


def run_burst(server, tids, queries):
    tickets = [server.submit(tids[k], s, n) for k, s, n in queries]
    return [server.wait(t) for t in tickets]


def check_same(x, y, kind):
    assert x.status == 0, x.error
    assert y.status == 0, y.error
    assert x.rows == y.rows
    if kind == 2:
        assert x.nulls == y.nulls
        assert hugeints(x.columns[0]) == hugeints(y.columns[0])
        np.testing.assert_array_equal(x.columns[1], y.columns[1])
        return
    for ca, cb, t in zip(x.columns, y.columns, x.types):
        if t == capi.FLOAT64 and kind == 1:
            np.testing.assert_allclose(ca, cb, rtol=1e-12, atol=0)
        else:
            np.testing.assert_array_equal(ca, cb)


def check_oracle(r, kind, start, size, h):
    sl = slice(start, start + size)
    if kind == 0:
        o1, o2 = oracle.filter_project_cfg0(h["a"][sl], h["b"][sl], h["c"][sl])
        assert r.rows == len(o1)
        np.testing.assert_array_equal(r.columns[0], o1)
        np.testing.assert_array_equal(r.columns[1], o2)
    elif kind == 1:
        if size == 0:
            assert r.rows == 0
            return
        ref = oracle.group_by_i64(h["key"][sl], h["v1"][sl], h["v2"][sl])
        np.testing.assert_array_equal(r.columns[0], ref["key"])
        np.testing.assert_array_equal(r.columns[1], ref["count"])
        assert hugeints(r.columns[2]) == ref["sum_v1"]
        np.testing.assert_array_equal(r.columns[3], ref["min_v1"])
        np.testing.assert_array_equal(r.columns[4], ref["max_v1"])
        np.testing.assert_allclose(r.columns[5], ref["sum_v2"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(r.columns[6], ref["avg_v2"], rtol=1e-12, atol=0)
    elif kind == 3:
        ox, oy = oracle.cross_filter_pairs(h["x"], h["y"][sl], 1 << 30)
        assert r.rows == len(ox)
        np.testing.assert_array_equal(r.columns[0], ox)
        np.testing.assert_array_equal(r.columns[1], oy)
    else:
        s, cnt = oracle.join_sum_count(h["bkey"], h["payload"], h["fk"][sl])
        assert int(r.columns[1][0]) == cnt
        assert r.nulls[0] == (cnt == 0)
        if cnt:
            assert hugeints(r.columns[0])[0] == s


def test_served_results_same_alone_and_in_the_mix(tables):
    h, d = tables
    queries = make_queries(len(h["a"]), 80, seed=1)
    with capi.Server(slots=4, batching=True) as mixed:
        tids = add_templates(mixed, d)
        in_mix = run_burst(mixed, tids, queries)
        stats = mixed.stats()
    with capi.Server(slots=1, batching=False) as alone_server:
        tids = add_templates(alone_server, d)
        alone = run_burst(alone_server, tids, queries)
    assert stats["max_batch"] > 1, stats  # the burst was batched
    assert max(r.batch_size for r in in_mix) > 1
    assert all(r.batch_size == 1 for r in alone)
    for (kind, start, size), x, y in zip(queries, in_mix, alone):
        check_same(x, y, kind)
        check_oracle(x, kind, start, size, h)


def test_served_open_loop_and_resubmission(tables):
    h, d = tables
    rng = np.random.default_rng(9)
    n = 90
    kinds = rng.integers(0, 3, n)
    sizes = np.exp(rng.uniform(np.log(10_000), np.log(1_000_000), n)).astype(np.uint64)
    starts = np.array([rng.integers(0, len(h["a"]) - s + 1) for s in sizes], dtype=np.uint64)
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, n))
    with capi.Server(slots=4, batching=True) as srv:
        tids = add_templates(srv, d)
        tmpl = np.array([tids[k] for k in kinds], dtype=np.int32)
        lat, bs, elapsed = srv.run_open_loop(tmpl, starts, sizes, arrivals)
        assert (lat > 0).all() and elapsed >= arrivals[-1]
        assert (bs >= 1).all()
        # the server keeps serving after the loop: a query resubmitted alone gives the in-mix answer
        r = srv.wait(srv.submit(tids[1], int(starts[0]), int(sizes[0])))
        check_oracle(r, 1, int(starts[0]), int(sizes[0]), h)


def test_served_admission_by_footprint(tables):
    h, d = tables
    # 96 MiB budget: a whole-table GROUP BY (2M rows) cannot fit and is rejected; small queries still run,
    # and batches stay within the budget.
    with capi.Server(slots=2, batching=True, memory_budget=96 << 20) as srv:
        tids = add_templates(srv, d)
        big = srv.submit(tids[1], 0, len(h["a"]))
        small = [srv.submit(tids[1], 1000 * i, 50_000) for i in range(12)]
        rb = srv.wait(big)
        assert rb.status != 0 and "budget" in rb.error
        for i, t in enumerate(small):
            r = srv.wait(t)
            check_oracle(r, 1, 1000 * i, 50_000, h)
        st = srv.stats()
        assert st["rejected"] == 1 and st["peak_reserved"] <= st["budget"]


def test_served_bad_arguments(tables):
    h, d = tables
    with capi.Server(slots=1) as srv:
        tids = add_templates(srv, d)
        with pytest.raises(capi.VgpuError):
            srv.submit(tids[0], len(h["a"]), 10)  # past the table
        with pytest.raises(capi.VgpuError):
            srv.submit(99, 0, 10)  # unknown template
        with pytest.raises(capi.VgpuError):
            srv.wait(12345)  # unknown ticket
        import torch
        with pytest.raises(capi.VgpuError):  # float GROUP BY key
            srv.add_group_by([d["v2"], d["v1"]], [], 0, [1], GB_AGGS[:1])
        with pytest.raises(capi.VgpuError):  # join key types differ
            srv.add_join_agg(d["a"], d["bkey"], d["payload"])
        del torch


def _join_reference(bkey, payload, fk):
    """Exact SUM(payload), COUNT(*) of fk JOIN build ON fk = bkey (Python integers)."""
    uk, inv = np.unique(bkey, return_inverse=True)
    cnt = np.bincount(inv, minlength=len(uk)).astype(np.int64)
    sums = [0] * len(uk)
    for i, p in zip(inv.tolist(), payload.tolist()):
        sums[i] += int(p)
    pos = np.searchsorted(uk, fk)
    pos_c = np.minimum(pos, len(uk) - 1)
    hit = (pos < len(uk)) & (uk[pos_c] == fk)
    total = sum(sums[j] for j in pos_c[hit].tolist())
    return total, int(cnt[pos_c[hit]].sum())


@pytest.mark.parametrize("case", ["sparse_keys", "wide_payload"])
def test_served_join_index_fallbacks(dev, case):
    """Join templates whose build side is not direct-addressable: keys spread far wider than their count
    (binary-search probe), or per-key payload sums past the packed entry (2^50-sized payloads)."""
    import torch
    rng = np.random.default_rng(17)
    nb, n = 50_000, 300_000
    if case == "sparse_keys":
        bkey = (rng.permutation(nb).astype(np.int64) * 1009 - 5_000_000)
        bkey = np.concatenate([bkey, bkey[:3000]])  # duplicates
        payload = rng.integers(-1000, 1000, len(bkey)).astype(np.int32)
        fk = np.where(rng.random(n) < 0.8, bkey[rng.integers(0, len(bkey), n)],
                      rng.integers(-10 ** 8, 10 ** 8, n)).astype(np.int64)
    else:
        bkey = np.concatenate([rng.permutation(nb), rng.integers(0, 100, 500)]).astype(np.int64)
        payload = rng.integers(-(1 << 50), 1 << 50, len(bkey)).astype(np.int64)
        fk = rng.integers(-10, nb + 1000, n).astype(np.int64)
    d = {k: torch.from_numpy(v).to(dev) for k, v in dict(bkey=bkey, payload=payload, fk=fk).items()}
    queries = [(int(s), int(c)) for s, c in zip(rng.integers(0, n // 2, 30), rng.integers(0, n // 2, 30))]
    queries += [(0, n), (5, 0), (n - 1, 1)]
    out = {}
    # batched; plain (both sides sorted per query); plain with the resident index, one query per probe
    for label, opts in (("batched", dict(batching=True)), ("plain", dict(batching=False)),
                        ("resident", dict(batching=False, resident_join=True))):
        with capi.Server(slots=3, **opts) as srv:
            t = srv.add_join_agg(d["fk"], d["bkey"], d["payload"])
            tickets = [srv.submit(t, s, c) for s, c in queries]
            out[label] = [srv.wait(k) for k in tickets]
    assert all(r.batch_size == 1 for r in out["resident"])
    for (s, c), x, y, z in zip(queries, out["batched"], out["plain"], out["resident"]):
        check_same(x, y, 2)
        check_same(x, z, 2)
        ref_sum, ref_cnt = _join_reference(bkey, payload, fk[s:s + c])
        assert int(x.columns[1][0]) == ref_cnt
        assert x.nulls[0] == (ref_cnt == 0)
        if ref_cnt:
            assert hugeints(x.columns[0])[0] == ref_sum
