"""Opt-in DuckDB semantics through GpuEngine::ExecutePlan (SetGpuExtendedSemanticsEnabled): NULL-aware
GROUP BY and join, FLOAT/DOUBLE group and join keys, composite and mixed-width join keys, MIN/MAX over
floating columns, and the keyless aggregate over no rows.

The reference operators refuse all of these (cuda_engine/src/operators/gpu_groupby.cu:24-33,
gpu_hash_join.cu:25-38), so there is no reference output to compare with; the expected results are DuckDB's
documented semantics restated in plain Python below, over small inputs:
  * GROUP BY: all NULL keys of a column form one group; -0.0 = +0.0 and NaN = NaN (NaN sorts above +inf);
    COUNT(*) counts rows, COUNT(v) non-NULL values; SUM/MIN/MAX/AVG skip NULLs and are NULL when a group has
    no non-NULL value; a keyless aggregate over no rows is one row (COUNT 0, the rest NULL).
  * INNER join: a NULL key (in any key column) matches nothing; tuples compare column by column; integers
    compare by value across widths; floating keys under the same equality as GROUP BY.
With the switch off, the same plans are refused exactly as the reference refuses them. Groups are compared
as maps (key tuple -> aggregates), join outputs as multisets of rows; exact for integers, HUGEINT, MIN/MAX;
SUM/AVG of doubles within 1e-12 relative (math.fsum as the exact sum)."""
import math
from collections import Counter, defaultdict

import numpy as np
import pytest

from vectisdb_amd.engine import (GpuAggregate, GpuAggregateKind as K, GpuEngine, GpuExpr, GpuOpType as Op,
                                 GpuPlanNode, GpuTableRef, GpuValueType as T, hugeint_to_int, make_column,
                                 set_gpu_extended_semantics_enabled, is_gpu_extended_semantics_enabled)

pytestmark = pytest.mark.gpu


@pytest.fixture
def extended():
    set_gpu_extended_semantics_enabled(True)
    yield
    set_gpu_extended_semantics_enabled(False)


def run(plan, inputs):
    return GpuEngine.instance().execute_plan(plan, inputs)


def scan(cols, table="t"):
    return GpuPlanNode(Op.SCAN, table=GpuTableRef("memory", "main", table), column_names=list(cols))


def group_by(keys, aggs, child):
    return GpuPlanNode(Op.GROUP_BY_AGGREGATE, children=[child], group_keys=list(keys), aggregates=list(aggs))


def join(lks, rks, outs, left, right):
    return GpuPlanNode(Op.HASH_JOIN, children=[left, right], build_keys=list(lks), probe_keys=list(rks),
                       output_columns=list(outs))


def values_of(col):
    """Python values of a result column, None where NULL."""
    mask = col.valid_mask()
    if col.type == T.HUGEINT:
        vals = hugeint_to_int(col.data[: col.row_count])
    else:
        vals = col.data[: col.row_count].tolist()
    return [v if m else None for v, m in zip(vals, mask)]


def canon(x):
    """DuckDB's grouping identity of a float: -0.0 -> 0.0, every NaN -> one NaN marker."""
    if isinstance(x, float):
        if math.isnan(x):
            return "NaN"
        if x == 0.0:
            return 0.0
    return x


def duck_min(vals, largest=False):
    """DuckDB's MIN/MAX over non-NULL values (NaN above everything)."""
    if not vals:
        return None
    nan = [v for v in vals if isinstance(v, float) and math.isnan(v)]
    rest = [v for v in vals if not (isinstance(v, float) and math.isnan(v))]
    if largest:
        return nan[0] if nan else max(rest)
    return min(rest) if rest else nan[0]


def same_value(got, want, rel=0.0):
    if want is None or got is None:
        return got is None and want is None
    if isinstance(want, float) and math.isnan(want):
        return isinstance(got, float) and math.isnan(got)
    if rel and isinstance(want, float):
        return abs(got - want) <= rel * abs(want) or got == want
    return got == want


def expected_groups(keys, vals, aggs, valid_keys, valid_vals):
    """keys: list of key arrays; vals: dict name -> array; aggs: (kind, name)."""
    n = len(keys[0]) if keys else len(next(iter(vals.values())))
    groups = defaultdict(list)
    for r in range(n):
        kt = tuple(canon(k[r].item()) if valid_keys[i][r] else None for i, k in enumerate(keys))
        groups[kt].append(r)
    out = {}
    for kt, rows in groups.items():
        res = []
        for kind, name in aggs:
            if kind == K.COUNT_STAR:
                res.append(len(rows))
                continue
            vs = [vals[name][r].item() for r in rows if valid_vals[name][r]]
            if kind == K.COUNT:
                res.append(len(vs))
            elif kind == K.SUM:
                res.append(None if not vs else (math.fsum(vs) if isinstance(vs[0], float) else sum(vs)))
            elif kind == K.AVG:
                res.append(None if not vs else (math.fsum(vs) if isinstance(vs[0], float) else sum(vs)) / len(vs))
            elif kind == K.MIN:
                res.append(duck_min(vs))
            else:
                res.append(duck_min(vs, largest=True))
        out[kt] = res
    return out


def check_groups(r, key_names, agg_specs, want):
    assert r.success, r.error_message
    nk = len(key_names)
    got_keys = [values_of(r.column(k)) for k in key_names]
    got_aggs = [values_of(r.columns[nk + i]) for i in range(len(agg_specs))]
    ng = r.columns[0].row_count
    assert ng == len(want), (ng, len(want))
    seen = set()
    for g in range(ng):
        kt = tuple(canon(got_keys[i][g]) for i in range(nk))
        assert kt in want, kt
        assert kt not in seen
        seen.add(kt)
        for i, (kind, _) in enumerate(agg_specs):
            rel = 1e-12 if kind in (K.SUM, K.AVG) else 0.0
            assert same_value(got_aggs[i][g], want[kt][i], rel), (kt, kind, got_aggs[i][g], want[kt][i])


AGG_SPECS = [(K.COUNT_STAR, ""), (K.COUNT, "v1"), (K.SUM, "v1"), (K.MIN, "v1"), (K.MAX, "v1"), (K.AVG, "v1"),
             (K.SUM, "v2"), (K.AVG, "v2"), (K.MIN, "v2"), (K.MAX, "v2"), (K.COUNT, "v2")]
OUT_TYPE = {(K.COUNT_STAR, ""): T.INT64, (K.COUNT, "v1"): T.INT64, (K.SUM, "v1"): T.HUGEINT,
            (K.MIN, "v1"): T.INT64, (K.MAX, "v1"): T.INT64, (K.AVG, "v1"): T.FLOAT64, (K.SUM, "v2"): T.FLOAT64,
            (K.AVG, "v2"): T.FLOAT64, (K.MIN, "v2"): T.FLOAT64, (K.MAX, "v2"): T.FLOAT64, (K.COUNT, "v2"): T.INT64}


def aggs_of(specs):
    return [GpuAggregate(kind, name, OUT_TYPE[(kind, name)]) for kind, name in specs]


def test_switch_defaults_off_and_refusals_stay(dev):
    assert not is_gpu_extended_semantics_enabled()
    valid = np.arange(256) != 0
    r = run(group_by(["k"], [GpuAggregate(K.COUNT_STAR, "", T.INT64)], scan(["k"])),
            [make_column("k", T.INT32, np.full(256, 5, np.int32), valid)])
    assert not r.success and "NULL" in r.error_message
    r = run(group_by(["k"], [GpuAggregate(K.COUNT_STAR, "", T.INT64)], scan(["k"])),
            [make_column("k", T.FLOAT64, np.full(64, 1.5))])
    assert not r.success and "integer/boolean" in r.error_message


@pytest.mark.parametrize("n,groups", [(5000, 40), (120_000, 3000)])
def test_group_by_nullable_keys_and_values(dev, extended, n, groups):
    rng = np.random.default_rng(n)
    k = rng.integers(-groups, groups, n)
    v1 = rng.integers(-10 ** 12, 10 ** 12, n)
    v2 = rng.standard_normal(n)
    vk = rng.random(n) > 0.1
    vv1 = rng.random(n) > 0.3
    vv2 = rng.random(n) > 0.3
    # a group whose values are all NULL
    k[:50] = 10 ** 9
    vv1[:50] = False
    vv2[:50] = False
    vk[:50] = True
    cols = [make_column("k", T.INT64, k, vk), make_column("v1", T.INT64, v1, vv1), make_column("v2", T.FLOAT64, v2, vv2)]
    r = run(group_by(["k"], aggs_of(AGG_SPECS), scan(["k", "v1", "v2"])), cols)
    want = expected_groups([k], {"v1": v1, "v2": v2}, AGG_SPECS, [vk], {"v1": vv1, "v2": vv2})
    check_groups(r, ["k"], AGG_SPECS, want)


def test_group_by_two_nullable_keys_int32_and_bool(dev, extended):
    rng = np.random.default_rng(5)
    n = 30_000
    a = rng.integers(-7, 7, n).astype(np.int32)
    b = rng.integers(0, 2, n).astype(np.uint8)
    va, vb = rng.random(n) > 0.2, rng.random(n) > 0.2
    v1 = rng.integers(-1000, 1000, n)
    specs = [(K.COUNT_STAR, ""), (K.SUM, "v1"), (K.MIN, "v1")]
    cols = [make_column("a", T.INT32, a, va), make_column("b", T.BOOLEAN, b, vb), make_column("v1", T.INT64, v1)]
    r = run(group_by(["a", "b"], aggs_of(specs), scan(["a", "b", "v1"])), cols)
    want = expected_groups([a, b], {"v1": v1}, specs, [va, vb], {"v1": np.ones(n, bool)})
    check_groups(r, ["a", "b"], specs, want)


@pytest.mark.parametrize("vtype,dtype", [(T.FLOAT64, np.float64), (T.FLOAT32, np.float32)])
def test_group_by_float_keys_duckdb_equality(dev, extended, vtype, dtype):
    rng = np.random.default_rng(9)
    n = 20_000
    base = np.array([0.0, -0.0, 1.5, -1.5, np.inf, -np.inf, np.nan, 2.25, -0.0, 1e-30], dtype)
    k = base[rng.integers(0, len(base), n)]
    # NaNs with other payloads and the sign bit set are the same group
    bits = k.view(np.uint64 if dtype == np.float64 else np.uint32).copy()
    nan_rows = np.isnan(k)
    if dtype == np.float64:
        bits[nan_rows] = np.where(rng.random(nan_rows.sum()) < 0.5, 0xFFF8000000000001, 0x7FF0000000000123)
    else:
        bits[nan_rows] = np.where(rng.random(nan_rows.sum()) < 0.5, 0xFFC00001, 0x7F800123).astype(np.uint32)
    k = bits.view(dtype)
    vk = rng.random(n) > 0.05
    v1 = rng.integers(-100, 100, n)
    v2 = rng.standard_normal(n)
    v2[::97] = np.nan
    specs = [(K.COUNT_STAR, ""), (K.SUM, "v1"), (K.MIN, "v2"), (K.MAX, "v2"), (K.SUM, "v2")]
    cols = [make_column("k", vtype, k, vk), make_column("v1", T.INT64, v1), make_column("v2", T.FLOAT64, v2)]
    r = run(group_by(["k"], aggs_of(specs), scan(["k", "v1", "v2"])), cols)
    assert r.success, r.error_message
    kv = values_of(r.column("k"))
    # one zero group, output as +0.0; one NaN group; NULL last
    zeros = [x for x in kv if x is not None and not math.isnan(x) and x == 0.0]
    assert len(zeros) == 1 and math.copysign(1.0, zeros[0]) == 1.0
    assert sum(1 for x in kv if x is not None and math.isnan(x)) == 1
    assert kv[-1] is None
    non_null = [x for x in kv if x is not None]
    assert all(canon(a) != canon(b) for a, b in zip(non_null, non_null[1:]))
    want = expected_groups([k], {"v1": v1, "v2": v2}, specs, [vk], {"v1": np.ones(n, bool), "v2": np.ones(n, bool)})
    check_groups(r, ["k"], specs, want)


def test_keyless_aggregate_over_empty_and_nullable_input(dev, extended):
    f = GpuPlanNode(Op.FILTER, children=[scan(["v"])],
                    expressions=[GpuExpr("((bool*)out)[idx] = (((const double*)inputs[0])[row]) > 99.0;", ["v"],
                                         T.BOOLEAN)])
    specs = [GpuAggregate(K.COUNT_STAR, "", T.INT64), GpuAggregate(K.SUM, "v", T.FLOAT64),
             GpuAggregate(K.COUNT, "v", T.INT64), GpuAggregate(K.MIN, "v", T.FLOAT64)]
    r = run(group_by([], specs, f), [make_column("v", T.FLOAT64, np.ones(4096))])
    assert r.success, r.error_message
    assert [values_of(c) for c in r.columns] == [[0], [None], [0], [None]]
    rng = np.random.default_rng(2)
    v = rng.standard_normal(10_000)
    valid = rng.random(10_000) > 0.5
    r = run(group_by([], specs, scan(["v"])), [make_column("v", T.FLOAT64, v, valid)])
    assert r.success, r.error_message
    got = [values_of(c)[0] for c in r.columns]
    assert got[0] == 10_000 and got[2] == int(valid.sum())
    assert abs(got[1] - math.fsum(v[valid])) <= 1e-12 * abs(math.fsum(v[valid]))
    assert got[3] == v[valid].min()


def expected_join(lkeys, rkeys, lvalid, rvalid, lpay, rpay):
    """Multiset of (left payload, right payload) over matching key tuples (NULL keys never match)."""
    index = defaultdict(list)
    for j in range(len(rkeys[0])):
        if all(v[j] for v in rvalid):
            index[tuple(canon(k[j].item()) for k in rkeys)].append(j)
    out = Counter()
    for i in range(len(lkeys[0])):
        if not all(v[i] for v in lvalid):
            continue
        for j in index.get(tuple(canon(k[i].item()) for k in lkeys), []):
            out[(lpay[i], rpay[j])] += 1
    return out


def join_rows(r, names):
    cols = [values_of(r.column(nm)) for nm in names]
    return Counter(zip(*cols))


def test_join_nullable_keys_and_payload(dev, extended):
    rng = np.random.default_rng(3)
    nl, nr = 4000, 9000
    lk = rng.integers(0, 3000, nl)
    rk = rng.integers(0, 3000, nr)
    lv, rv = rng.random(nl) > 0.15, rng.random(nr) > 0.15
    lp = rng.integers(-50, 50, nl).astype(np.int32)
    lpv = rng.random(nl) > 0.3
    rp = rng.standard_normal(nr)
    cols = [make_column("lk", T.INT64, lk, lv), make_column("lp", T.INT32, lp, lpv),
            make_column("rk", T.INT64, rk, rv), make_column("rp", T.FLOAT64, rp)]
    r = run(join(["lk"], ["rk"], ["lp", "rp"], scan(["lk", "lp"], "l"), scan(["rk", "rp"], "r")), cols)
    assert r.success, r.error_message
    lpay = [int(x) if v else None for x, v in zip(lp, lpv)]
    want = expected_join([lk], [rk], [lv], [rv], lpay, rp.tolist())
    assert join_rows(r, ["lp", "rp"]) == want


def test_join_composite_and_mixed_width_keys(dev, extended):
    rng = np.random.default_rng(4)
    nl, nr = 3000, 5000
    la = rng.integers(-20, 20, nl).astype(np.int32)
    lb = rng.integers(-(1 << 40), -(1 << 40) + 30, nl)
    ra = rng.integers(-20, 20, nr)                       # INT64 against the left's INT32
    rb = rng.integers(-(1 << 40), -(1 << 40) + 30, nr)
    rbv = rng.random(nr) > 0.1
    lid = np.arange(nl, dtype=np.int64)
    rid = np.arange(nr, dtype=np.int64)
    cols = [make_column("la", T.INT32, la), make_column("lb", T.INT64, lb), make_column("lid", T.INT64, lid),
            make_column("ra", T.INT64, ra), make_column("rb", T.INT64, rb, rbv), make_column("rid", T.INT64, rid)]
    r = run(join(["la", "lb"], ["ra", "rb"], ["lid", "rid"], scan(["la", "lb", "lid"], "l"),
                 scan(["ra", "rb", "rid"], "r")), cols)
    assert r.success, r.error_message
    want = expected_join([la, lb], [ra, rb], [np.ones(nl, bool), np.ones(nl, bool)], [np.ones(nr, bool), rbv],
                         lid.tolist(), rid.tolist())
    assert sum(want.values()) > 1000
    assert join_rows(r, ["lid", "rid"]) == want


def test_join_float_keys(dev, extended):
    lk = np.array([0.0, -0.0, np.nan, 1.5, -np.inf, 2.0], np.float64)
    rk = np.array([-0.0, np.nan, 1.5, 1.5, -np.inf, 3.0, 0.0], np.float64)
    rk_bits = rk.view(np.uint64).copy()
    rk_bits[1] = 0xFFF8000000000007  # another NaN
    rk = rk_bits.view(np.float64)
    lid, rid = np.arange(len(lk), dtype=np.int64), np.arange(len(rk), dtype=np.int64)
    cols = [make_column("lk", T.FLOAT64, lk), make_column("lid", T.INT64, lid), make_column("rk", T.FLOAT64, rk),
            make_column("rid", T.INT64, rid)]
    r = run(join(["lk"], ["rk"], ["lid", "rid"], scan(["lk", "lid"], "l"), scan(["rk", "rid"], "r")), cols)
    assert r.success, r.error_message
    ones = [np.ones(len(lk), bool)]
    want = expected_join([lk], [rk], ones, [np.ones(len(rk), bool)], lid.tolist(), rid.tolist())
    assert join_rows(r, ["lid", "rid"]) == want
    assert want[(0, 0)] == 1 and want[(1, 6)] == 1 and want[(2, 1)] == 1


def test_join_aggregate_with_nullable_payload(dev, extended):
    """A keyless aggregate over the join with a nullable input leaves the fused join-aggregate path."""
    rng = np.random.default_rng(8)
    nb, np_ = 2000, 20_000
    bk = rng.permutation(nb).astype(np.int64)
    pay = rng.integers(-1000, 1000, nb).astype(np.int64)
    pv = rng.random(nb) > 0.25
    fk = rng.integers(0, 2 * nb, np_)
    j = join(["fk"], ["bk"], ["pay"], scan(["fk"], "p"), scan(["bk", "pay"], "b"))
    specs = [GpuAggregate(K.SUM, "pay", T.HUGEINT), GpuAggregate(K.COUNT, "pay", T.INT64),
             GpuAggregate(K.COUNT_STAR, "", T.INT64), GpuAggregate(K.MAX, "pay", T.INT64)]
    r = run(group_by([], specs, j),
            [make_column("fk", T.INT64, fk), make_column("bk", T.INT64, bk), make_column("pay", T.INT64, pay, pv)])
    assert r.success, r.error_message
    pos = {int(k): i for i, k in enumerate(bk)}
    hits = [pos[int(x)] for x in fk if int(x) in pos]
    vals = [int(pay[i]) for i in hits if pv[i]]
    assert [values_of(c)[0] for c in r.columns] == [sum(vals), len(vals), len(hits), max(vals)]
    # an empty join: one row of NULLs with COUNT 0
    r = run(group_by([], specs, j), [make_column("fk", T.INT64, fk + 10 * nb), make_column("bk", T.INT64, bk),
                                      make_column("pay", T.INT64, pay, pv)])
    assert r.success, r.error_message
    assert [values_of(c)[0] for c in r.columns] == [None, 0, 0, None]


def test_cross_product_over_nulls_still_refused(dev, extended):
    cp = GpuPlanNode(Op.CROSS_PRODUCT, children=[scan(["x"], "l"), scan(["y"], "r")], output_columns=["x", "y"])
    r = run(cp, [make_column("x", T.INT32, np.arange(8, dtype=np.int32), np.arange(8) != 3),
                 make_column("y", T.INT32, np.arange(8, dtype=np.int32))])
    assert not r.success and "NULL" in r.error_message
