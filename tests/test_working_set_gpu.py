"""EstimateWorkingSet against the device memory a plan really takes: the device pool's peak bytes in use while
GpuEngine::ExecutePlan runs the plan from host columns (every plan temporary of this engine comes from that
pool). The bar is the reference's own: predicted / measured peak within [0.40, 4.0]
(cuda_engine/tests/test_gpu_cost_model.cpp:37-40, :77-130), here for the filter/projection, GROUP BY,
materialized join, join + aggregate and cross product + filter plans, with the row counts and widths the
extension would annotate (estimated_rows, estimated_row_width_bytes)."""
import numpy as np
import pytest

from vectisdb_amd import capi
from vectisdb_amd.engine import (GpuAggregate, GpuAggregateKind as K, GpuEngine, GpuExpr, GpuOpType as Op,
                                 GpuPlanNode, GpuTableRef, GpuValueType as T, estimate_working_set, make_column)
from vectisdb_amd.workloads import CFG0_PRED_A, CFG0_PRED_C, CFG0_PROJ_1, CFG0_PROJ_2

pytestmark = pytest.mark.gpu

LO, HI = 0.40, 4.0


def scan(cols, rows, width, table="t"):
    return GpuPlanNode(Op.SCAN, table=GpuTableRef("memory", "main", table), column_names=list(cols),
                       estimated_rows=rows, estimated_row_width_bytes=width)


def measured_peak(plan, inputs):
    capi.pool_stats()
    before = capi.pool_stats()["in_use"]
    capi.pool_reset_peak()
    r = GpuEngine.instance().execute_plan(plan, inputs)
    assert r.success, r.error_message
    return capi.pool_stats()["peak"] - before, r


def check(plan, inputs, label):
    est = estimate_working_set(plan, 0).total_bytes
    peak, r = measured_peak(plan, inputs)
    assert peak > 0, label
    ratio = est / peak
    assert LO <= ratio <= HI, f"{label}: estimate {est} B, measured peak {peak} B, ratio {ratio:.3f}"
    return r


def test_filter_projection_config0_shape(dev):
    n = 4_000_000
    rng = np.random.default_rng(1)
    a = rng.integers(0, 1_000_000, n).astype(np.int32)
    b = rng.integers(-(1 << 40), 1 << 40, n)
    c = rng.random(n)
    kept = int(((a < 500000) & (c > 0.25)).sum())
    s = scan(["a", "b", "c"], n, 20)
    f = GpuPlanNode(Op.FILTER, children=[s], estimated_rows=kept, estimated_row_width_bytes=20,
                    expressions=[GpuExpr(CFG0_PRED_A, ["a"], T.BOOLEAN), GpuExpr(CFG0_PRED_C, ["c"], T.BOOLEAN)])
    p = GpuPlanNode(Op.PROJECTION, children=[f], estimated_rows=kept, estimated_row_width_bytes=16,
                    expressions=[GpuExpr(CFG0_PROJ_1, ["a", "b"], T.INT64), GpuExpr(CFG0_PROJ_2, ["c", "a"], T.FLOAT64)])
    r = check(p, [make_column("a", T.INT32, a), make_column("b", T.INT64, b), make_column("c", T.FLOAT64, c)],
              "filter + projection")
    assert r.columns[0].row_count == kept


@pytest.mark.parametrize("groups", [64, 200_000])
def test_group_by_seven_aggregates(dev, groups):
    n = 4_000_000
    rng = np.random.default_rng(groups)
    key = rng.integers(0, groups, n)
    v1 = rng.integers(-1_000_000, 1_000_001, n)
    v2 = rng.standard_normal(n)
    s = scan(["key", "v1", "v2"], n, 24)
    aggs = [GpuAggregate(K.COUNT_STAR, "", T.INT64), GpuAggregate(K.SUM, "v1", T.HUGEINT),
            GpuAggregate(K.MIN, "v1", T.INT64), GpuAggregate(K.MAX, "v1", T.INT64),
            GpuAggregate(K.SUM, "v2", T.FLOAT64), GpuAggregate(K.AVG, "v2", T.FLOAT64)]
    g = GpuPlanNode(Op.GROUP_BY_AGGREGATE, children=[s], group_keys=["key"], aggregates=aggs,
                    estimated_rows=groups, estimated_row_width_bytes=64)
    check(g, [make_column("key", T.INT64, key), make_column("v1", T.INT64, v1), make_column("v2", T.FLOAT64, v2)],
          f"GROUP BY, {groups} groups")


def test_materialized_join(dev):
    nl, nr = 1_000_000, 2_000_000
    rng = np.random.default_rng(3)
    lk = rng.permutation(nl).astype(np.int64)
    rk = rng.integers(0, 2 * nl, nr)
    matches = int((rk < nl).sum())
    left = scan(["lk", "a"], nl, 16, "l")
    right = scan(["rk", "b"], nr, 16, "r")
    j = GpuPlanNode(Op.HASH_JOIN, children=[left, right], build_keys=["lk"], probe_keys=["rk"],
                    output_columns=["a", "b"], estimated_rows=matches, estimated_row_width_bytes=16)
    r = check(j, [make_column("lk", T.INT64, lk), make_column("a", T.INT64, np.arange(nl)),
                  make_column("rk", T.INT64, rk), make_column("b", T.INT64, np.arange(nr))], "materialized join")
    assert r.columns[0].row_count == matches


@pytest.mark.parametrize("form", ["table", "sort_merge"])
def test_join_aggregate(dev, form):
    nb, np_ = 1 << 20, 1 << 23
    rng = np.random.default_rng(4)
    bk = rng.permutation(nb).astype(np.int64)
    pay = rng.integers(0, 1000, nb).astype(np.int32)
    fk = rng.integers(0, 2 * nb, np_)
    build = scan(["key", "payload"], nb, 12, "b")
    probe = scan(["fk"], np_, 8, "p")
    matches = int((fk < nb).sum())
    j = GpuPlanNode(Op.HASH_JOIN, children=[probe, build], build_keys=["fk"], probe_keys=["key"],
                    output_columns=["payload"], estimated_rows=matches, estimated_row_width_bytes=4)
    # SUM/COUNT runs as a per-key table probed by the larger side; MAX needs the sort-merge form
    first = GpuAggregate(K.SUM, "payload", T.HUGEINT) if form == "table" else GpuAggregate(K.MAX, "payload", T.INT32)
    g = GpuPlanNode(Op.GROUP_BY_AGGREGATE, children=[j], aggregates=[first, GpuAggregate(K.COUNT_STAR, "", T.INT64)],
                    estimated_rows=1, estimated_row_width_bytes=24)
    check(g, [make_column("fk", T.INT64, fk), make_column("key", T.INT64, bk), make_column("payload", T.INT32, pay)],
          f"join + aggregate ({form})")


def test_cross_product_filter(dev):
    nl, nr = 2048, 8192
    rng = np.random.default_rng(5)
    x = rng.integers(0, 1 << 30, nl).astype(np.int32)
    y = rng.integers(0, 1 << 30, nr).astype(np.int32)
    kept = int((x[:, None].astype(np.int64) + y[None, :] < (1 << 30)).sum())
    cp = GpuPlanNode(Op.CROSS_PRODUCT, children=[scan(["x"], nl, 4, "l"), scan(["y"], nr, 4, "r")],
                     output_columns=["x", "y"], estimated_rows=nl * nr, estimated_row_width_bytes=8)
    pred = "((bool*)out)[idx] = (((const int*)inputs[0])[row] + ((const int*)inputs[1])[row]) < 1073741824;"
    f = GpuPlanNode(Op.FILTER, children=[cp], expressions=[GpuExpr(pred, ["x", "y"], T.BOOLEAN)],
                    estimated_rows=kept, estimated_row_width_bytes=8)
    r = check(f, [make_column("x", T.INT32, x), make_column("y", T.INT32, y)], "cross product + filter")
    assert r.columns[0].row_count == kept
