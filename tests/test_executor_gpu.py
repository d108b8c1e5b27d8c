"""GpuEngine::ExecutePlan through the C ABI: every case of the reference's executor suite
(cuda_engine/tests/test_gpu_executor.cpp:122-1440, same inputs, same assertions), plus the paths this
engine adds (fused filter/projection/cross kernels, join-aggregate without materialization, HUGEINT SUM,
column pruning, wide expressions, the column cache). Expected values are computed independently with numpy
on the CPU from the same inputs."""
import numpy as np
import pytest

from vectisdb_amd.engine import (GpuAggregate, GpuAggregateKind as K, GpuEngine, GpuExecuteOptions, GpuExpr,
                                 GpuOpType as Op, GpuPlanNode, GpuTableRef, GpuValueType as T, GpuColumnCache,
                                 hugeint_to_int, make_column)

pytestmark = pytest.mark.gpu


def run(plan, inputs, options=None):
    return GpuEngine.instance().execute_plan(plan, inputs, options)


def scan(cols, table="t"):
    return GpuPlanNode(Op.SCAN, table=GpuTableRef("memory", "main", table), column_names=list(cols))


def expr(body, inputs, out):
    return GpuExpr(body, list(inputs), out)


def filt(child, *preds):
    return GpuPlanNode(Op.FILTER, children=[child], expressions=list(preds))


def proj(child, *exprs):
    return GpuPlanNode(Op.PROJECTION, children=[child], expressions=list(exprs))


def group_by(keys, aggs, child):
    return GpuPlanNode(Op.GROUP_BY_AGGREGATE, children=[child], group_keys=list(keys), aggregates=list(aggs))


def agg(kind, col, out):
    return GpuAggregate(kind, col, out)


def join(lk, rk, outs, left, right):
    return GpuPlanNode(Op.HASH_JOIN, children=[left, right], build_keys=[lk], probe_keys=[rk],
                       output_columns=list(outs))


def cross(outs, left, right):
    return GpuPlanNode(Op.CROSS_PRODUCT, children=[left, right], output_columns=list(outs))


def ok(result, ncols):
    assert result.success, result.error_message
    assert result.error_message == ""
    assert len(result.columns) == ncols
    return result


def group_map(keys, *vals):
    return {tuple(k) if isinstance(k, tuple) else k: tuple(v[i] for v in vals) for i, k in enumerate(keys)}


# ---------------------------------------------------------------- FILTER / PROJECTION


def test_projection_runs_on_gpu(dev):
    n = 50000
    revenue = (np.arange(n) % 997).astype(np.float64) + 0.5
    p = proj(scan(["revenue"]), expr("((double*)out)[idx] = (((const double*)inputs[0])[row]) * 1.18;",
                                     ["revenue"], T.FLOAT64))
    r = ok(run(p, [make_column("revenue", T.FLOAT64, revenue)]), 1)
    assert r.columns[0].name == "col0" and r.columns[0].row_count == n
    np.testing.assert_array_equal(r.columns[0].data, revenue * 1.18)  # one IEEE multiply: bit-exact


def test_filter_compacts_on_gpu(dev):
    idv = (np.arange(20000) % 500).astype(np.int32)
    f = filt(scan(["id"]), expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) > 100;", ["id"], T.BOOLEAN))
    r = ok(run(f, [make_column("id", T.INT32, idv)]), 1)
    np.testing.assert_array_equal(r.columns[0].data, idv[idv > 100])


def test_filter_then_projection_pipeline(dev):
    qty = (np.arange(12000) % 10).astype(np.int32)
    f = filt(scan(["qty"]), expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) > 3;", ["qty"], T.BOOLEAN))
    p = proj(f, expr("((int*)out)[idx] = (((const int*)inputs[0])[row]) * 2;", ["qty"], T.INT32))
    r = ok(run(p, [make_column("qty", T.INT32, qty)]), 1)
    np.testing.assert_array_equal(r.columns[0].data, qty[qty > 3] * 2)


def test_multi_column_filter_gathers_every_column(dev):
    n = 5000
    a = (np.arange(n) % 7).astype(np.int32)
    b = np.arange(n, dtype=np.float64)
    f = filt(scan(["a", "b"]), expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) > 4;", ["a"], T.BOOLEAN))
    r = ok(run(f, [make_column("a", T.INT32, a), make_column("b", T.FLOAT64, b)]), 2)
    assert [c.name for c in r.columns] == ["a", "b"]
    np.testing.assert_array_equal(r.columns[0].data, a[a > 4])
    np.testing.assert_array_equal(r.columns[1].data, b[a > 4])


def test_filter_matching_nothing_produces_empty_result(dev):
    f = filt(scan(["id"]), expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) > 999;", ["id"], T.BOOLEAN))
    r = ok(run(f, [make_column("id", T.INT32, np.ones(4096, np.int32))]), 1)
    assert r.columns[0].row_count == 0


def test_nullable_column_passes_through_bare_scan(dev):
    n = 256
    idv = np.arange(n, dtype=np.int32)
    valid = np.arange(n) != 0
    r = ok(run(scan(["id"]), [make_column("id", T.INT32, idv, valid)]), 1)
    assert r.columns[0].row_count == n
    np.testing.assert_array_equal(r.columns[0].valid_mask(), valid)
    np.testing.assert_array_equal(r.columns[0].data, idv)


def test_nullable_column_propagates_through_projection(dev):
    n = 128
    v = np.arange(n, dtype=np.int32)
    valid = ~np.isin(np.arange(n), [3, 100])
    p = proj(scan(["v"]), expr("((int*)out)[idx] = (((const int*)inputs[0])[row]) + 1;", ["v"], T.INT32))
    r = ok(run(p, [make_column("v", T.INT32, v, valid)]), 1)
    got_valid = r.columns[0].valid_mask()
    np.testing.assert_array_equal(got_valid, valid)
    np.testing.assert_array_equal(r.columns[0].data[got_valid], (v + 1)[valid])


def test_nullable_column_excluded_by_filter_predicate(dev):
    n = 64
    idv = np.arange(n, dtype=np.int32)
    v = np.arange(n, dtype=np.int32)
    valid = np.arange(n) != 10
    f = filt(scan(["id", "v"]), expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) > 5;", ["v"], T.BOOLEAN))
    r = ok(run(f, [make_column("id", T.INT32, idv), make_column("v", T.INT32, v, valid)]), 2)
    keep = (v > 5) & valid
    np.testing.assert_array_equal(r.columns[0].data, idv[keep])
    # the carried nullable column keeps its validity for the surviving rows
    np.testing.assert_array_equal(r.columns[1].valid_mask(), valid[keep])


def test_malformed_plan_fails_cleanly(dev):
    j = GpuPlanNode(Op.HASH_JOIN, children=[scan([])])
    r = run(j, [])
    assert not r.success and r.error_message


def test_repeated_execution_hits_kernel_cache(dev):
    n = 8192
    v = np.arange(n, dtype=np.float64)
    p = proj(scan(["v"]), expr("((double*)out)[idx] = (((const double*)inputs[0])[row]) + 7.0;", ["v"],
                               T.FLOAT64))
    for _ in range(3):
        r = ok(run(p, [make_column("v", T.FLOAT64, v)]), 1)
        np.testing.assert_array_equal(r.columns[0].data, v + 7.0)


# ---------------------------------------------------------------- GROUP BY


def test_group_by_sum_on_gpu(dev):
    n = 100000
    k = (np.arange(n) % 1000).astype(np.int32)
    v = (np.arange(n) % 64).astype(np.float64)
    r = ok(run(group_by(["k"], [agg(K.SUM, "v", T.FLOAT64)], scan(["k", "v"])),
               [make_column("k", T.INT32, k), make_column("v", T.FLOAT64, v)]), 2)
    keys, sums = r.columns[0].data, r.columns[1].data
    assert len(set(keys.tolist())) == len(keys) == 1000
    expect = np.bincount(k, weights=v, minlength=1000)
    np.testing.assert_array_equal(sums, expect[keys])  # small integers: exact in any order
    assert r.columns[1].name == "agg_0" and r.columns[0].name == "k"


def test_group_by_count_star_on_gpu(dev):
    n = 60000
    k = ((np.arange(n, dtype=np.int64) ** 2) % 251).astype(np.int64)
    r = ok(run(group_by(["k"], [agg(K.COUNT_STAR, "", T.INT64)], scan(["k"])), [make_column("k", T.INT64, k)]), 2)
    keys, counts = r.columns[0].data, r.columns[1].data
    uniq, cnt = np.unique(k, return_counts=True)
    assert len(keys) == len(uniq) and len(set(keys.tolist())) == len(keys)
    assert dict(zip(keys.tolist(), counts.tolist())) == dict(zip(uniq.tolist(), cnt.tolist()))
    assert counts.sum() == n


def test_group_by_min_max_on_gpu(dev):
    n = 50000
    i = np.arange(n, dtype=np.int64)
    k = (i % 97).astype(np.int32)
    v = (i * 7919) % 100000 - 50000
    r = ok(run(group_by(["k"], [agg(K.MIN, "v", T.INT64), agg(K.MAX, "v", T.INT64)], scan(["k", "v"])),
               [make_column("k", T.INT32, k), make_column("v", T.INT64, v)]), 3)
    keys = r.columns[0].data
    for j, key in enumerate(keys):
        sel = v[k == key]
        assert r.columns[1].data[j] == sel.min() and r.columns[2].data[j] == sel.max()
    assert len(keys) == 97


def test_group_by_multiple_keys_on_gpu(dev):
    n = 80000
    i = np.arange(n)
    k1 = (i % 13 - 6).astype(np.int32)
    k2 = (i % 7 - 3).astype(np.int64)
    v = (i % 32).astype(np.float64)
    r = ok(run(group_by(["k1", "k2"], [agg(K.SUM, "v", T.FLOAT64)], scan(["k1", "k2", "v"])),
               [make_column("k1", T.INT32, k1), make_column("k2", T.INT64, k2), make_column("v", T.FLOAT64, v)]), 3)
    expect = {}
    for a_, b_, c_ in zip(k1.tolist(), k2.tolist(), v.tolist()):
        expect[(a_, b_)] = expect.get((a_, b_), 0.0) + c_
    got = {(a_, b_): s for a_, b_, s in zip(r.columns[0].data.tolist(), r.columns[1].data.tolist(),
                                            r.columns[2].data.tolist())}
    assert len(got) == len(r.columns[0].data) == 91
    assert got == expect


def test_group_by_global_aggregate_on_gpu(dev):
    n = 30000
    v = (np.arange(n) % 128).astype(np.float64)
    r = ok(run(group_by([], [agg(K.COUNT_STAR, "", T.INT64), agg(K.SUM, "v", T.FLOAT64)], scan(["v"])),
               [make_column("v", T.FLOAT64, v)]), 2)
    assert r.columns[0].data.tolist() == [n]
    assert r.columns[1].data.tolist() == [v.sum()]


def test_group_by_after_filter_on_gpu(dev):
    n = 40000
    k = (np.arange(n) % 50).astype(np.int32)
    v = (np.arange(n) % 30).astype(np.int32)
    f = filt(scan(["k", "v"]), expr("((bool*)out)[idx] = (((const int*)inputs[1])[row]) > 10;", ["k", "v"],
                                    T.BOOLEAN))
    r = ok(run(group_by(["k"], [agg(K.COUNT_STAR, "", T.INT64)], f),
               [make_column("k", T.INT32, k), make_column("v", T.INT32, v)]), 2)
    uniq, cnt = np.unique(k[v > 10], return_counts=True)
    assert dict(zip(r.columns[0].data.tolist(), r.columns[1].data.tolist())) == dict(zip(uniq.tolist(), cnt.tolist()))


def test_group_by_sum_of_float_promotes_to_double(dev):
    n = 20000
    k = (np.arange(n) % 4).astype(np.int32)
    v = (np.arange(n) % 256).astype(np.float32)
    r = ok(run(group_by(["k"], [agg(K.SUM, "v", T.FLOAT64)], scan(["k", "v"])),
               [make_column("k", T.INT32, k), make_column("v", T.FLOAT32, v)]), 2)
    assert r.columns[1].type == T.FLOAT64 and len(r.columns[0].data) == 4
    expect = np.bincount(k, weights=v.astype(np.float64))
    np.testing.assert_array_equal(r.columns[1].data, expect[r.columns[0].data])


def test_group_by_high_cardinality_on_gpu(dev):
    groups = 200000
    n = groups * 2
    k = (np.arange(n) % groups).astype(np.int64)
    v = np.arange(n, dtype=np.int64)
    r = ok(run(group_by(["k"], [agg(K.COUNT_STAR, "", T.INT64), agg(K.MAX, "v", T.INT64)], scan(["k", "v"])),
               [make_column("k", T.INT64, k), make_column("v", T.INT64, v)]), 3)
    keys = r.columns[0].data
    assert len(keys) == groups and len(np.unique(keys)) == groups
    assert (r.columns[1].data == 2).all()
    np.testing.assert_array_equal(r.columns[2].data, keys + groups)


def test_group_by_empty_input_produces_no_groups(dev):
    f = filt(scan(["k"]), expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) > 999;", ["k"], T.BOOLEAN))
    r = ok(run(group_by(["k"], [agg(K.COUNT_STAR, "", T.INT64)], f),
               [make_column("k", T.INT32, np.full(8192, 3, np.int32))]), 2)
    assert r.columns[0].row_count == 0 and r.columns[1].row_count == 0


def test_group_by_float_key_refused_cleanly(dev):
    r = run(group_by(["k"], [agg(K.COUNT_STAR, "", T.INT64)], scan(["k"])),
            [make_column("k", T.FLOAT64, np.full(1024, 1.5))])
    assert not r.success and "integer/boolean" in r.error_message


def test_nullable_column_refused_for_group_by_aggregate(dev):
    valid = np.arange(256) != 0
    r = run(group_by(["k"], [agg(K.COUNT_STAR, "", T.INT64)], scan(["k"])),
            [make_column("k", T.INT32, np.full(256, 5, np.int32), valid)])
    assert not r.success and "NULL" in r.error_message


def test_group_by_float_min_max_refused_cleanly(dev):
    r = run(group_by(["k"], [agg(K.MAX, "v", T.FLOAT64)], scan(["k", "v"])),
            [make_column("k", T.INT32, np.ones(1024, np.int32)), make_column("v", T.FLOAT64, np.full(1024, 2.0))])
    assert not r.success and r.error_message


def test_group_by_global_aggregate_over_empty_input_refused_cleanly(dev):
    f = filt(scan(["v"]), expr("((bool*)out)[idx] = (((const double*)inputs[0])[row]) > 99.0;", ["v"], T.BOOLEAN))
    r = run(group_by([], [agg(K.SUM, "v", T.FLOAT64)], f), [make_column("v", T.FLOAT64, np.ones(4096))])
    assert not r.success and "NULL" in r.error_message


def test_group_by_with_projected_smallint_key(dev):
    n = 30000
    k = (np.arange(n) % 7).astype(np.int16)
    p = proj(scan(["k"]), expr("((short*)out)[idx] = ((const short*)inputs[0])[row];", ["k"], T.INT16))
    r = ok(run(group_by(["col0"], [agg(K.COUNT_STAR, "", T.INT64)], p), [make_column("k", T.INT16, k)]), 2)
    assert r.columns[0].type == T.INT16 and r.columns[1].type == T.INT64
    uniq, cnt = np.unique(k, return_counts=True)
    assert dict(zip(r.columns[0].data.tolist(), r.columns[1].data.tolist())) == dict(zip(uniq.tolist(), cnt.tolist()))


def test_group_by_with_projected_computed_key(dev):
    n = 50000
    k = np.arange(n, dtype=np.int32)
    v = (np.arange(n) % 100).astype(np.int32)
    p = proj(scan(["k", "v"]),
             expr("((int*)out)[idx] = (((const int*)inputs[0])[row]) % 10;", ["k"], T.INT32),
             expr("((double*)out)[idx] = (double)(((const int*)inputs[0])[row]);", ["v"], T.FLOAT64))
    r = ok(run(group_by(["col0"], [agg(K.SUM, "col1", T.FLOAT64)], p),
               [make_column("k", T.INT32, k), make_column("v", T.INT32, v)]), 2)
    expect = np.bincount(k % 10, weights=v.astype(np.float64))
    assert len(r.columns[0].data) == 10
    np.testing.assert_array_equal(r.columns[1].data, expect[r.columns[0].data])


def test_group_by_integer_sum_is_hugeint(dev):
    """SUM over BIGINT returns DuckDB's HUGEINT, exactly, including sums beyond int64."""
    rng = np.random.default_rng(7)
    n = 300000
    k = rng.integers(0, 50, n).astype(np.int64)
    v = rng.integers(-(2 ** 62), 2 ** 62, n, dtype=np.int64)
    r = ok(run(group_by(["k"], [agg(K.SUM, "v", T.HUGEINT), agg(K.AVG, "v", T.FLOAT64),
                                agg(K.COUNT, "v", T.INT64)], scan(["k", "v"])),
               [make_column("k", T.INT64, k), make_column("v", T.INT64, v)]), 4)
    sums = hugeint_to_int(r.columns[1].data)
    for j, key in enumerate(r.columns[0].data.tolist()):
        exact = sum(int(x) for x in v[k == key])
        assert sums[j] == exact
        cnt = int((k == key).sum())
        assert r.columns[3].data[j] == cnt
        assert abs(r.columns[2].data[j] - exact / cnt) <= 1e-12 * abs(exact / cnt) + 1e-300


def test_group_by_integer_sum_to_double_refused(dev):
    r = run(group_by(["k"], [agg(K.SUM, "v", T.FLOAT64)], scan(["k", "v"])),
            [make_column("k", T.INT32, np.ones(16, np.int32)), make_column("v", T.INT32, np.ones(16, np.int32))])
    assert not r.success and "SUM" in r.error_message


# ---------------------------------------------------------------- HASH_JOIN


def join_inputs(lid, lv, rid, rv, kt=T.INT32):
    return [make_column("lid", kt, lid), make_column("lv", T.INT32, lv),
            make_column("rid", kt, rid), make_column("rv", T.INT32, rv)]


def sorted_pairs(a, b):
    return sorted(zip(a.tolist(), b.tolist()))


def test_inner_join_on_gpu(dev):
    lid = np.arange(10000, dtype=np.int32)
    rid = np.arange(5000, dtype=np.int32)
    r = ok(run(join("lid", "rid", ["lv", "rv"], scan(["lid", "lv"]), scan(["rid", "rv"])),
               join_inputs(lid, lid * 2, rid, rid * 3)), 2)
    assert r.columns[0].row_count == 5000
    assert sorted_pairs(r.columns[0].data, r.columns[1].data) == sorted((i * 2, i * 3) for i in range(5000))


def test_join_many_to_many_on_gpu(dev):
    lid, lv, rid, rv = [], [], [], []
    for kk in range(50):
        for i in range(3):
            lid.append(kk)
            lv.append(kk * 10 + i)
        for j in range(4):
            rid.append(kk)
            rv.append(kk * 100 + j)
    a = lambda x: np.array(x, dtype=np.int32)
    r = ok(run(join("lid", "rid", ["lv", "rv"], scan(["lid", "lv"]), scan(["rid", "rv"])),
               join_inputs(a(lid), a(lv), a(rid), a(rv))), 2)
    expect = sorted((kk * 10 + i, kk * 100 + j) for kk in range(50) for i in range(3) for j in range(4))
    assert r.columns[0].row_count == 600
    assert sorted_pairs(r.columns[0].data, r.columns[1].data) == expect


def test_join_with_negative_keys_on_gpu(dev):
    n = 2000
    lid = np.arange(n, dtype=np.int64) - n
    r = ok(run(join("lid", "rid", ["lv", "rv"], scan(["lid", "lv"]), scan(["rid", "rv"])),
               join_inputs(lid, np.arange(n, dtype=np.int32), lid.copy(), np.arange(n, dtype=np.int32) + 1000,
                           kt=T.INT64)), 2)
    assert sorted_pairs(r.columns[0].data, r.columns[1].data) == [(i, i + 1000) for i in range(n)]


def test_join_output_column_order_is_respected(dev):
    n = 1000
    i = np.arange(n, dtype=np.int32)
    r = ok(run(join("lid", "rid", ["rv", "lv"], scan(["lid", "lv"]), scan(["rid", "rv"])),
               join_inputs(i, i + 7, i.copy(), i + 9)), 2)
    assert [c.name for c in r.columns] == ["rv", "lv"]
    assert r.columns[0].row_count == n
    np.testing.assert_array_equal(r.columns[0].data, r.columns[1].data + 2)


def test_join_after_filter_on_gpu(dev):
    n = 4000
    i = np.arange(n, dtype=np.int32)
    lv = (i % 100).astype(np.int32)
    f = filt(scan(["lid", "lv"]), expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) > 50;", ["lv"],
                                       T.BOOLEAN))
    r = ok(run(join("lid", "rid", ["lv", "rv"], f, scan(["rid", "rv"])), join_inputs(i, lv, i.copy(), i.copy())), 2)
    assert r.columns[0].row_count == int((lv > 50).sum())


def test_join_matching_nothing_produces_empty_result(dev):
    n = 2048
    r = ok(run(join("lid", "rid", ["lv", "rv"], scan(["lid", "lv"]), scan(["rid", "rv"])),
               join_inputs(np.ones(n, np.int32), np.full(n, 5, np.int32), np.full(n, 999, np.int32),
                           np.full(n, 7, np.int32))), 2)
    assert r.columns[0].row_count == 0 and r.columns[1].row_count == 0


def test_join_mismatched_key_types_refused_cleanly(dev):
    r = run(join("lid", "rid", ["lid", "rid"], scan(["lid"]), scan(["rid"])),
            [make_column("lid", T.INT32, np.full(512, 3, np.int32)), make_column("rid", T.INT64, np.full(512, 3))])
    assert not r.success and "same type" in r.error_message


def test_join_float_key_refused_cleanly(dev):
    r = run(join("lid", "rid", ["lid", "rid"], scan(["lid"]), scan(["rid"])),
            [make_column("lid", T.FLOAT64, np.full(512, 1.5)), make_column("rid", T.FLOAT64, np.full(512, 1.5))])
    assert not r.success and "integer/boolean" in r.error_message


def test_join_ambiguous_output_column_refused_cleanly(dev):
    r = run(join("id", "id", ["id"], scan(["id"]), scan(["id"])),
            [make_column("id", T.INT32, np.ones(512, np.int32)), make_column("id", T.INT32, np.ones(512, np.int32))])
    assert not r.success and "ambiguous" in r.error_message


def test_join_composite_key_refused_cleanly(dev):
    node = join("l1", "r1", ["l1"], scan(["l1"]), scan(["r1"]))
    node.build_keys.append("l1")
    node.probe_keys.append("r1")
    r = run(node, [make_column("l1", T.INT32, np.ones(512, np.int32)), make_column("r1", T.INT32, np.ones(512, np.int32))])
    assert not r.success and "exactly one equality key" in r.error_message


def test_global_aggregate_over_join_without_materializing(dev):
    """SELECT SUM(payload), COUNT(*), AVG(payload), MIN(fk), MAX(fk) FROM probe JOIN build ON fk = key:
    GROUP BY without keys directly over the join runs as a join-aggregate; compared with numpy."""
    rng = np.random.default_rng(11)
    nb, npr = 1 << 14, 1 << 17
    key = rng.permutation(nb).astype(np.int64)
    payload = rng.integers(0, 1000, nb).astype(np.int32)
    fk = np.where(rng.random(npr) < 0.9, rng.integers(0, nb, npr), rng.integers(nb, 2 * nb, npr)).astype(np.int64)
    j = join("key", "fk", ["payload", "fk"], scan(["key", "payload"], "build"), scan(["fk"], "probe"))
    g = group_by([], [agg(K.SUM, "payload", T.HUGEINT), agg(K.COUNT_STAR, "", T.INT64),
                      agg(K.AVG, "payload", T.FLOAT64), agg(K.MIN, "fk", T.INT64), agg(K.MAX, "fk", T.INT64)], j)
    r = ok(run(g, [make_column("key", T.INT64, key), make_column("payload", T.INT32, payload),
                   make_column("fk", T.INT64, fk)]), 5)
    inv = np.empty(nb, np.int64)
    inv[key] = np.arange(nb)
    hit = fk < nb
    matched = payload[inv[fk[hit]]].astype(np.int64)
    assert hugeint_to_int(r.columns[0].data) == [int(matched.sum())]
    assert r.columns[1].data.tolist() == [int(hit.sum())]
    assert r.columns[2].data[0] == pytest.approx(matched.sum() / hit.sum(), rel=1e-12)
    assert r.columns[3].data.tolist() == [int(fk[hit].min())]
    assert r.columns[4].data.tolist() == [int(fk[hit].max())]
    assert [c.name for c in r.columns] == [f"agg_{i}" for i in range(5)]


def test_global_aggregate_over_empty_join_refused(dev):
    j = join("lid", "rid", ["lv"], scan(["lid", "lv"]), scan(["rid", "rv"]))
    g = group_by([], [agg(K.SUM, "lv", T.HUGEINT)], j)
    r = run(g, join_inputs(np.ones(64, np.int32), np.ones(64, np.int32), np.full(64, 9, np.int32),
                           np.ones(64, np.int32)))
    assert not r.success and "NULL" in r.error_message


def test_group_by_keys_over_join(dev):
    """A keyed GROUP BY over a join materializes the join, then groups it."""
    lid = np.arange(100, dtype=np.int32)
    lv = (lid % 5).astype(np.int32)
    rid = np.repeat(np.arange(100, dtype=np.int32), 3)
    rv = np.arange(300, dtype=np.int32)
    j = join("lid", "rid", ["lv", "rv"], scan(["lid", "lv"]), scan(["rid", "rv"]))
    r = ok(run(group_by(["lv"], [agg(K.SUM, "rv", T.HUGEINT)], j), join_inputs(lid, lv, rid, rv)), 2)
    expect = {g_: int(rv[(lv[rid] == g_)].sum()) for g_ in range(5)}
    assert dict(zip(r.columns[0].data.tolist(), hugeint_to_int(r.columns[1].data))) == expect


# ---------------------------------------------------------------- CROSS_PRODUCT


def test_cross_product_on_gpu(dev):
    r = ok(run(cross(["lv", "rv"], scan(["lv"]), scan(["rv"])),
               [make_column("lv", T.INT32, [10, 20, 30]), make_column("rv", T.INT32, [1, 2, 3, 4])]), 2)
    assert r.columns[0].data.tolist() == [10] * 4 + [20] * 4 + [30] * 4
    assert r.columns[1].data.tolist() == [1, 2, 3, 4] * 3


def test_cross_product_mixed_widths_and_column_order(dev):
    r = ok(run(cross(["rk", "lval", "lk"], scan(["lk", "lval"]), scan(["rk"])),
               [make_column("lk", T.INT32, [7, 8]), make_column("lval", T.FLOAT64, [1.5, 2.5]),
                make_column("rk", T.INT64, [100, 200, 300])]), 3)
    assert r.columns[0].data.tolist() == [100, 200, 300] * 2
    assert r.columns[1].data.tolist() == [1.5] * 3 + [2.5] * 3
    assert r.columns[2].data.tolist() == [7] * 3 + [8] * 3


def test_projection_over_cross_product_on_gpu(dev):
    lval = np.arange(500) * 0.25
    rval = np.arange(400) * 0.5 + 1.0
    p = proj(cross(["lval", "rval"], scan(["lval"]), scan(["rval"])),
             expr("((double*)out)[idx] = (((const double*)inputs[0])[row]) * (((const double*)inputs[1])[row]);",
                  ["lval", "rval"], T.FLOAT64))
    r = ok(run(p, [make_column("lval", T.FLOAT64, lval), make_column("rval", T.FLOAT64, rval)]), 1)
    np.testing.assert_array_equal(r.columns[0].data, np.outer(lval, rval).ravel())


def test_cross_product_empty_side_produces_no_rows(dev):
    r = ok(run(cross(["lv", "rv"], scan(["lv"]), scan(["rv"])),
               [make_column("lv", T.INT32, [1, 2, 3]), make_column("rv", T.INT32, np.zeros(0, np.int32))]), 2)
    assert [c.row_count for c in r.columns] == [0, 0]
    assert [c.type for c in r.columns] == [T.INT32, T.INT32]


def test_cross_product_oversize_refused_cleanly(dev):
    n = 70000
    r = run(cross(["lv", "rv"], scan(["lv"]), scan(["rv"])),
            [make_column("lv", T.INT32, np.ones(n, np.int32)), make_column("rv", T.INT32, np.full(n, 2, np.int32))])
    assert not r.success and "too large" in r.error_message


def test_cross_product_ambiguous_output_column_refused_cleanly(dev):
    r = run(cross(["id"], scan(["id"]), scan(["id"])),
            [make_column("id", T.INT32, [1, 2]), make_column("id", T.INT32, [3, 4])])
    assert not r.success and "ambiguous" in r.error_message


def test_filter_over_cross_product_fused(dev):
    """SELECT x, y FROM L, R WHERE x + y < C: the predicate runs inside the pair enumeration; output in
    the reference's left-major order."""
    rng = np.random.default_rng(3)
    x = rng.integers(0, 1 << 30, 300).astype(np.int32)
    y = rng.integers(0, 1 << 30, 4096).astype(np.int32)
    f = filt(cross(["x", "y"], scan(["x"], "L"), scan(["y"], "R")),
             expr("((bool*)out)[idx] = ((long long)((const int*)inputs[0])[row] + (long long)((const int*)inputs[1])[row]) < 1073741824LL;",
                  ["x", "y"], T.BOOLEAN))
    r = ok(run(f, [make_column("x", T.INT32, x), make_column("y", T.INT32, y)]), 2)
    xx, yy = np.repeat(x, len(y)), np.tile(y, len(x))
    keep = xx.astype(np.int64) + yy < (1 << 30)
    np.testing.assert_array_equal(r.columns[0].data, xx[keep])
    np.testing.assert_array_equal(r.columns[1].data, yy[keep])


def test_projection_over_filter_over_cross_product(dev):
    x = np.arange(37, dtype=np.int32)
    y = np.arange(53, dtype=np.int32) * 3
    f = filt(cross(["x", "y"], scan(["x"]), scan(["y"])),
             expr("((bool*)out)[idx] = (((const int*)inputs[0])[row] % 3) != (((const int*)inputs[1])[row] % 5);",
                  ["x", "y"], T.BOOLEAN))
    p = proj(f, expr("((long long*)out)[idx] = (long long)((const int*)inputs[0])[row] * 1000 + ((const int*)inputs[1])[row];",
                     ["x", "y"], T.INT64))
    r = ok(run(p, [make_column("x", T.INT32, x), make_column("y", T.INT32, y)]), 1)
    xx, yy = np.repeat(x, len(y)).astype(np.int64), np.tile(y, len(x)).astype(np.int64)
    keep = (xx % 3) != (yy % 5)
    np.testing.assert_array_equal(r.columns[0].data, (xx * 1000 + yy)[keep])


# ---------------------------------------------------------------- engine paths beyond the reference suite


def test_filter_projection_config_query(dev):
    """SELECT a*2+b, c*1.5+a FROM t WHERE a < 500000 AND c > 0.25 (two conjuncts, mixed types),
    bit-exact against numpy's IEEE evaluation of the same operations."""
    rng = np.random.default_rng(5)
    n = 1 << 18
    a = rng.integers(0, 10 ** 6, n).astype(np.int32)
    b = rng.integers(-(2 ** 40), 2 ** 40, n).astype(np.int64)
    c = rng.random(n)
    f = filt(scan(["a", "b", "c"]),
             expr("((bool*)out)[idx] = (((const int*)inputs[0])[row]) < 500000;", ["a"], T.BOOLEAN),
             expr("((bool*)out)[idx] = (((const double*)inputs[0])[row]) > 0.25;", ["c"], T.BOOLEAN))
    p = proj(f, expr("((long long*)out)[idx] = (long long)(((const int*)inputs[0])[row]) * 2 + ((const long long*)inputs[1])[row];",
                     ["a", "b"], T.INT64),
             expr("((double*)out)[idx] = ((const double*)inputs[0])[row] * 1.5 + (double)(((const int*)inputs[1])[row]);",
                  ["c", "a"], T.FLOAT64))
    r = ok(run(p, [make_column("a", T.INT32, a), make_column("b", T.INT64, b), make_column("c", T.FLOAT64, c)]), 2)
    keep = (a < 500000) & (c > 0.25)
    np.testing.assert_array_equal(r.columns[0].data, a[keep].astype(np.int64) * 2 + b[keep])
    np.testing.assert_array_equal(r.columns[1].data, c[keep] * 1.5 + a[keep].astype(np.float64))


def test_filter_carrying_hugeint_column(dev):
    """A FILTER above a GROUP BY whose output holds a HUGEINT column (HAVING-shaped): the HUGEINT column is
    carried through the compaction unchanged."""
    k = (np.arange(1000) % 10).astype(np.int64)
    v = np.arange(1000, dtype=np.int64) * (1 << 40)
    g = group_by(["k"], [agg(K.SUM, "v", T.HUGEINT)], scan(["k", "v"]))
    f = filt(g, expr("((bool*)out)[idx] = (((const long long*)inputs[0])[row]) >= 5;", ["k"], T.BOOLEAN))
    r = ok(run(f, [make_column("k", T.INT64, k), make_column("v", T.INT64, v)]), 2)
    got = dict(zip(r.columns[0].data.tolist(), hugeint_to_int(r.columns[1].data)))
    assert got == {g_: int(v[k == g_].astype(object).sum()) for g_ in range(5, 10)}


def test_wide_projection_beyond_one_kernel(dev):
    """20 outputs over 18 columns with a filter: more than one kernel's argument tables."""
    rng = np.random.default_rng(9)
    n = 5000
    cols = [rng.integers(-1000, 1000, n).astype(np.int64) for _ in range(18)]
    inputs = [make_column(f"c{i}", T.INT64, c) for i, c in enumerate(cols)]
    f = filt(scan([f"c{i}" for i in range(18)]),
             expr("((bool*)out)[idx] = (((const long long*)inputs[0])[row]) > 0;", ["c0"], T.BOOLEAN))
    exprs = [expr("((long long*)out)[idx] = ((const long long*)inputs[0])[row] - ((const long long*)inputs[1])[row];",
                  [f"c{i % 18}", f"c{(i + 1) % 18}"], T.INT64) for i in range(20)]
    r = ok(run(proj(f, *exprs), inputs), 20)
    keep = cols[0] > 0
    for i in range(20):
        np.testing.assert_array_equal(r.columns[i].data, (cols[i % 18] - cols[(i + 1) % 18])[keep])


def test_predicate_must_be_boolean(dev):
    f = filt(scan(["a"]), expr("((int*)out)[idx] = 1;", ["a"], T.INT32))
    r = run(f, [make_column("a", T.INT32, np.arange(10, dtype=np.int32))])
    assert not r.success and "BOOLEAN" in r.error_message


def test_missing_column_reported(dev):
    p = proj(scan(["a"]), expr("((int*)out)[idx] = ((const int*)inputs[0])[row];", ["zz"], T.INT32))
    r = run(p, [make_column("a", T.INT32, np.arange(10, dtype=np.int32))])
    assert not r.success and "'zz'" in r.error_message


def test_too_few_inputs_reported(dev):
    r = run(scan(["a", "b"]), [make_column("a", T.INT32, np.arange(10, dtype=np.int32))])
    assert not r.success and "more scanned columns" in r.error_message


def test_cache_staging_then_chunk_replay_and_whole_scan(dev):
    """Stage two chunks of a table into the column cache, then replay chunk by chunk (row-independent
    plan) and serve the whole scan (GROUP BY) without host inputs."""
    ref = GpuTableRef("memory", "main", "cached_t")
    GpuColumnCache.invalidate_all()
    rng = np.random.default_rng(1)
    k = rng.integers(0, 20, 6000).astype(np.int64)
    v = rng.integers(-100, 100, 6000).astype(np.int64)
    chunks = [(0, 3500), (3500, 6000)]
    sc = GpuPlanNode(Op.SCAN, table=ref, column_names=["k", "v"])
    GpuColumnCache.begin_staging(ref)
    for lo, hi in chunks:
        ok(run(sc, [make_column("k", T.INT64, k[lo:hi]), make_column("v", T.INT64, v[lo:hi])],
               GpuExecuteOptions(stage_scan_to_cache=True)), 2)
    GpuColumnCache.commit_staging(ref, 6000)
    assert GpuColumnCache.replayable_chunk_rows(ref, ["k", "v"]) == [3500, 2500]
    assert GpuColumnCache.current_bytes() >= 6000 * 16
    p = proj(sc, expr("((long long*)out)[idx] = ((const long long*)inputs[0])[row] * 3 + ((const long long*)inputs[1])[row];",
                      ["k", "v"], T.INT64))
    for i, (lo, hi) in enumerate(chunks):
        r = ok(run(p, [], GpuExecuteOptions(cache_chunk_index=i)), 1)
        np.testing.assert_array_equal(r.columns[0].data, k[lo:hi] * 3 + v[lo:hi])
    g = group_by(["k"], [agg(K.SUM, "v", T.HUGEINT), agg(K.COUNT_STAR, "", T.INT64)], sc)
    r = ok(run(g, [], GpuExecuteOptions(serve_whole_scan_from_cache=True)), 3)
    got = dict(zip(r.columns[0].data.tolist(), zip(hugeint_to_int(r.columns[1].data), r.columns[2].data.tolist())))
    assert got == {int(kk): (int(v[k == kk].sum()), int((k == kk).sum())) for kk in np.unique(k)}
    bad = run(p, [make_column("k", T.INT64, k)], GpuExecuteOptions(cache_chunk_index=0))
    assert not bad.success and "no host inputs" in bad.error_message
    GpuColumnCache.invalidate_table(ref)
    gone = run(p, [], GpuExecuteOptions(cache_chunk_index=0))
    assert not gone.success and "retry the query" in gone.error_message


def test_cache_options_mutually_exclusive(dev):
    r = run(scan(["a"]), [], GpuExecuteOptions(stage_scan_to_cache=True, cache_chunk_index=0))
    assert not r.success and "cannot be requested together" in r.error_message
    r = run(scan(["a"]), [], GpuExecuteOptions(cache_chunk_index=0, serve_whole_scan_from_cache=True))
    assert not r.success and "at once" in r.error_message


def test_large_upload_pipelined(dev):
    """A column above the 16 MiB staging piece goes through the alternating page-locked slots."""
    n = (40 << 20) // 8 + 123
    v = np.arange(n, dtype=np.int64) * 7 - 3
    p = proj(scan(["v"]), expr("((long long*)out)[idx] = ((const long long*)inputs[0])[row] + 1;", ["v"], T.INT64))
    r = ok(run(p, [make_column("v", T.INT64, v)]), 1)
    np.testing.assert_array_equal(r.columns[0].data, v + 1)
