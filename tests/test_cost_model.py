"""Routing cost model through the C ABI (no GPU): the reference's plan-shape and cost-model suites
(cuda_engine/tests/test_plan_shape.cpp, the CPU-only cases of test_gpu_cost_model.cpp) with the same
inputs and assertions, plus a randomized comparison against the reference's own gpu_cost_model.cpp compiled
from its sources into oracle/_ref (IsRowIndependent, AnalyzePlanShape, EstimateColumnsSize and
DetermineExecutionMode must agree exactly; EstimateWorkingSet models this engine's allocations and is
checked for its structural properties instead)."""
import ctypes as C
import random
from pathlib import Path

import pytest

from vectisdb_amd.engine import (ExecutionMode, GpuAggregate, GpuAggregateKind as K, GpuExpr, GpuOpType as Op,
                                 GpuPlanNode, GpuTableRef, GpuValueType as T, GpuWorkDensity as D,
                                 analyze_plan_shape, determine_execution_mode, estimate_columns_size,
                                 estimate_working_set, hash_source, is_row_independent, kOffloadOpsPerRow)

ADD, MUL, SQRT, TRANS = 1.0, 2.0, 4.0, 30.0
REF_LIB = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libvref.so"


def scan(cols, rows, width):
    return GpuPlanNode(Op.SCAN, table=GpuTableRef("memory", "main", "bench"), column_names=list(cols),
                       estimated_rows=rows, estimated_row_width_bytes=width)


def ex(ops):
    return GpuExpr("((double*)out)[idx] = 0.0;", [], T.FLOAT64, estimated_ops_per_row=ops)


def projection(costs, rows, scanned_columns):
    return GpuPlanNode(Op.PROJECTION, expressions=[ex(c) for c in costs],
                       children=[scan(["a", "b", "c"], rows, scanned_columns * 8)], estimated_rows=rows,
                       estimated_row_width_bytes=len(costs) * 8)


def cross_of(n, m):
    return GpuPlanNode(Op.CROSS_PRODUCT, children=[scan(["av"], n, 8), scan(["bv"], m, 8)], estimated_rows=n * m,
                       estimated_row_width_bytes=16)


# ---------------------------------------------------------------- plan shape (test_plan_shape.cpp)


def test_cheap_projection_is_declined_on_density():
    s = analyze_plan_shape(projection([ADD], 20000000, 2))
    assert s.scanned_rows == 20000000 and s.scanned_bytes == 20000000 * 16
    assert s.ops_per_scanned_row == 1.0 and s.density == D.INSUFFICIENT and s.amplification == 1.0


def test_heavy_math_projection_offloads_on_density_alone():
    heavy = 3 * TRANS + SQRT + TRANS + ADD + ADD + MUL + 4 * ADD
    s = analyze_plan_shape(projection([heavy], 20000000, 3))
    assert s.ops_per_scanned_row == heavy >= kOffloadOpsPerRow
    assert s.density == D.SUFFICIENT and s.amplification == 1.0 and s.ops_per_scanned_byte > 5.0


def test_moderately_dense_projections_are_inconclusive():
    b = analyze_plan_shape(projection([MUL + SQRT + ADD], 20000000, 3))
    assert b.ops_per_scanned_row == 7.0 and b.density == D.INCONCLUSIVE
    w = analyze_plan_shape(projection([MUL, ADD, SQRT, ADD], 20000000, 3))
    assert w.ops_per_scanned_row == 8.0 and w.density == D.INCONCLUSIVE


def test_filter_is_charged_over_its_input_rows():
    f = GpuPlanNode(Op.FILTER, expressions=[ex(ADD + ADD)], children=[scan(["a", "b"], 20000000, 16)],
                    estimated_rows=200000, estimated_row_width_bytes=16)
    s = analyze_plan_shape(f)
    assert s.total_ops == 2.0 * 20000000.0 and s.ops_per_scanned_row == 2.0


def test_cross_product_earns_density_through_expansion():
    c = cross_of(4472, 4472)
    p = GpuPlanNode(Op.PROJECTION, expressions=[ex(ADD)], children=[c], estimated_rows=c.estimated_rows,
                    estimated_row_width_bytes=8)
    s = analyze_plan_shape(p)
    assert s.scanned_rows == 8944 and s.ops_per_scanned_row > 2000.0 and s.density == D.SUFFICIENT


def test_plan_with_no_priced_expressions_is_unmeasured_rather_than_cheap():
    g = GpuPlanNode(Op.GROUP_BY_AGGREGATE, group_keys=["c"], children=[scan(["a", "c"], 20000000, 16)],
                    estimated_rows=1000, estimated_row_width_bytes=16)
    s = analyze_plan_shape(g)
    assert s.total_ops == 0.0 and s.ops_per_scanned_row == 0.0 and s.density == D.INCONCLUSIVE
    assert s.max_rows == 20000000 and s.amplification == 1.0


def test_aggregate_over_cross_product_survives_with_no_priced_expressions():
    a = GpuPlanNode(Op.GROUP_BY_AGGREGATE, children=[cross_of(4472, 4472)], estimated_rows=1,
                    estimated_row_width_bytes=8)
    s = analyze_plan_shape(a)
    assert s.total_ops == 0.0 and s.density == D.INCONCLUSIVE and s.amplification > 2000.0


def test_plan_that_scans_nothing_is_not_a_divide_by_zero():
    p = GpuPlanNode(Op.PROJECTION, expressions=[ex(TRANS)], estimated_rows=1000000, estimated_row_width_bytes=8)
    s = analyze_plan_shape(p)
    assert s.scanned_rows == 0 and s.scanned_bytes == 0
    assert s.ops_per_scanned_row == s.total_ops and s.ops_per_scanned_row == s.ops_per_scanned_row
    assert s.density == D.SUFFICIENT


def test_unpriced_expressions_do_not_talk_routing_into_offloading():
    s = analyze_plan_shape(projection([0.0, 0.0], 20000000, 2))
    assert s.total_ops == 0.0 and s.density != D.SUFFICIENT and s.amplification == 1.0


def test_auto_bypass_cache_threshold():
    budget = 3435973836
    p40 = projection([1.0], 40000000, 5)
    assert estimate_columns_size(p40) == 40000000 * 40
    assert determine_execution_mode(p40, budget, 0.75) == ExecutionMode.PIPELINE_STREAM_CACHED
    p150 = projection([1.0], 150000000, 5)
    assert estimate_columns_size(p150) == 150000000 * 40
    assert determine_execution_mode(p150, budget, 0.75) == ExecutionMode.PIPELINE_STREAM_UNCACHED
    assert determine_execution_mode(scan(["c"], 312500000, 8), budget, 0.75) == ExecutionMode.PIPELINE_STREAM_CACHED
    assert determine_execution_mode(scan(["c"], 350000000, 8), budget, 0.75) == ExecutionMode.PIPELINE_STREAM_UNCACHED


# ---------------------------------------------------------------- working set (CPU cases of test_gpu_cost_model.cpp)


def test_chunk_cap_applies_only_to_row_independent_plans():
    p = GpuPlanNode(Op.PROJECTION, expressions=[GpuExpr("x", ["v"], T.FLOAT64)], children=[scan(["v"], 50000000, 8)],
                    estimated_rows=50000000, estimated_row_width_bytes=8)
    uncapped, capped = estimate_working_set(p, 0), estimate_working_set(p, 2 * 1024 * 1024)
    assert not uncapped.chunk_capped and capped.chunk_capped
    assert capped.total_bytes < uncapped.total_bytes / 10
    g = GpuPlanNode(Op.GROUP_BY_AGGREGATE, group_keys=["k"], children=[scan(["k", "v"], 50000000, 12)],
                    estimated_rows=10, estimated_row_width_bytes=12)
    grouped = estimate_working_set(g, 2 * 1024 * 1024)
    assert not grouped.chunk_capped and not is_row_independent(g) and is_row_independent(p)
    # The sort's per-input-row buffers are invisible from the 10-row output estimate and must be charged.
    assert grouped.total_bytes > 2000 * 1024 * 1024


def test_oversized_cross_product_saturates_rather_than_wrapping():
    c = GpuPlanNode(Op.CROSS_PRODUCT, children=[scan(["a"], 100000, 8), scan(["b"], 100000, 8)],
                    estimated_rows=10 ** 10, estimated_row_width_bytes=16)
    e = estimate_working_set(c, 0)
    assert e.total_bytes >= 10 ** 10 * 16 and e.largest_operator == "CROSS_PRODUCT"
    huge = GpuPlanNode(Op.CROSS_PRODUCT, children=[scan(["a"], 1, 8), scan(["b"], 1, 8)],
                       estimated_rows=2 ** 64 - 1, estimated_row_width_bytes=16)
    assert estimate_working_set(huge, 0).total_bytes == 2 ** 64 - 1


def test_working_set_models_this_engines_operators():
    """Per-operator terms of this engine's allocations (fused FILTER writes only survivors of carried
    columns; GROUP BY charges sort double buffers and bound-sized outputs; a global aggregate needs no
    sort)."""
    n = 1000000
    f = GpuPlanNode(Op.FILTER, expressions=[ex(1.0)], children=[scan(["a"], n, 8)], estimated_rows=n // 2,
                    estimated_row_width_bytes=8)
    assert estimate_working_set(f, 0).total_bytes == n * 8 + (n * 8 + (n // 4096 + 1) * 8)
    g = GpuPlanNode(Op.GROUP_BY_AGGREGATE, group_keys=["k"], aggregates=[GpuAggregate(K.SUM, "v", T.HUGEINT)],
                    children=[scan(["k", "v"], n, 16)], estimated_rows=10, estimated_row_width_bytes=24)
    assert estimate_working_set(g, 0).total_bytes == n * 16 + n * (16 + 16) + n * (24 + 8)
    glob = GpuPlanNode(Op.GROUP_BY_AGGREGATE, aggregates=[GpuAggregate(K.SUM, "v", T.HUGEINT)],
                       children=[scan(["v"], n, 8)], estimated_rows=1, estimated_row_width_bytes=16)
    assert estimate_working_set(glob, 0).total_bytes == n * 8 + 16 + 8


# ---------------------------------------------------------------- exact agreement with the reference's code


def ref_lib():
    if not REF_LIB.exists():
        pytest.skip("oracle/_ref/libvref.so not built (needs the reference tree)")
    L = C.CDLL(str(REF_LIB))
    L.vref_analyze_plan_shape.argtypes = [C.c_char_p, C.POINTER(C.c_double)]
    L.vref_plan_shape_counts.argtypes = [C.c_char_p, C.POINTER(C.c_uint64)]
    L.vref_is_row_independent.argtypes = [C.c_char_p]
    L.vref_estimate_columns_size.argtypes = [C.c_char_p, C.POINTER(C.c_uint64)]
    L.vref_determine_execution_mode.argtypes = [C.c_char_p, C.c_uint64, C.c_double]
    L.vref_estimate_working_set.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_int),
                                            C.c_char_p, C.c_size_t]
    L.vref_hash_source.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
    return L


def random_plan(rng, depth=0):
    big = rng.random() < 0.15
    rows = rng.choice([0, 1, 7, 1000, 2 ** 20, 10 ** 8, 2 ** 40, 2 ** 63]) if big else rng.randrange(0, 10 ** 7)
    width = rng.choice([0, 4, 8, 12, 16, 40, 2 ** 31])
    if depth >= 3 or rng.random() < 0.3:
        return scan(["a", "b"][: rng.randint(0, 2)], rows, width)
    op = rng.choice([Op.FILTER, Op.PROJECTION, Op.GROUP_BY_AGGREGATE, Op.HASH_JOIN, Op.CROSS_PRODUCT])
    nkids = 2 if op in (Op.HASH_JOIN, Op.CROSS_PRODUCT) else 1
    node = GpuPlanNode(op, children=[random_plan(rng, depth + 1) for _ in range(nkids)], estimated_rows=rows,
                       estimated_row_width_bytes=width)
    if op in (Op.FILTER, Op.PROJECTION):
        node.expressions = [ex(rng.choice([0.0, 0.5, 1.0, 2.0, 7.5, 30.0, 1e6])) for _ in range(rng.randint(0, 3))]
    if op == Op.GROUP_BY_AGGREGATE:
        node.group_keys = ["a"] * rng.randint(0, 2)
        node.aggregates = [GpuAggregate(rng.choice(list(K)), "b", T.INT64) for _ in range(rng.randint(0, 3))]
    return node


def test_routing_functions_match_reference_code_exactly():
    L = ref_lib()
    rng = random.Random(1234)
    for _ in range(400):
        plan = random_plan(rng)
        js = plan.to_json().encode()
        out = (C.c_double * 8)()
        assert L.vref_analyze_plan_shape(js, out) == 0
        counts = (C.c_uint64 * 3)()
        assert L.vref_plan_shape_counts(js, counts) == 0
        mine = analyze_plan_shape(plan)
        assert (mine.scanned_rows, mine.scanned_bytes, mine.max_rows) == tuple(counts)
        for got, want in zip([mine.total_ops, mine.ops_per_scanned_row, mine.ops_per_scanned_byte,
                              float(int(mine.density)), mine.amplification], list(out)[3:]):
            assert got == want or (got != got and want != want)
        assert int(is_row_independent(plan)) == L.vref_is_row_independent(js)
        cs = C.c_uint64()
        assert L.vref_estimate_columns_size(js, C.byref(cs)) == 0
        assert estimate_columns_size(plan) == cs.value
        for budget in [0, 1, 2 ** 20, 3435973836, 2 ** 40, 2 ** 64 - 1]:
            for thr in [0.0, 0.5, 0.75, 1.0]:
                assert int(determine_execution_mode(plan, budget, thr)) == L.vref_determine_execution_mode(js, budget, thr)
        # Working set: this engine's own model, but the chunk-cap rule is the reference's.
        total, capped = C.c_uint64(), C.c_int()
        name = C.create_string_buffer(64)
        for cap in [0, 2 ** 21]:
            assert L.vref_estimate_working_set(js, cap, C.byref(total), C.byref(capped), name, 64) == 0
            assert int(estimate_working_set(plan, cap).chunk_capped) == capped.value


def test_hash_source_matches_reference_code():
    L = ref_lib()
    rng = random.Random(5)
    texts = ["", "abc", "a" * 55, "b" * 56, "c" * 64, "((double*)out)[idx] = 1.0; [opt:strict]"]
    texts += ["".join(chr(rng.randrange(1, 256)) for _ in range(rng.randrange(0, 300))) for _ in range(50)]
    for t in texts:
        raw = t.encode("latin-1")
        buf = C.create_string_buffer(65)
        L.vref_hash_source(raw, len(raw), buf)
        mine = C.create_string_buffer(65)
        from vectisdb_amd import capi
        capi.check(capi.load().vgpu_hash_source(raw, len(raw), mine))
        assert mine.value == buf.value
