// plan_json.h -- GpuPlanNode trees from JSON, for callers that cross a C boundary (ctypes, tests, tools).
//
// Header-only and written against the names of gpu_engine.hpp only, so any header that declares the
// vector_gpu plan types with those names can be used with it. Schema (every key optional unless noted):
//   {"op": "SCAN"|"FILTER"|"PROJECTION"|"HASH_JOIN"|"GROUP_BY_AGGREGATE"|"CROSS_PRODUCT"   (required),
//    "mode": 0..2, "estimated_rows": n, "estimated_row_width_bytes": n,
//    "table": {"catalog": s, "schema": s, "table": s}, "column_names": [s],
//    "expressions": [{"source": body, "hash": s, "inputs": [s], "output_type": T, "window": "RUNNING_SUM",
//                     "ops": x}],
//    "build_keys": [s], "probe_keys": [s], "output_columns": [s], "group_keys": [s],
//    "aggregate_exprs": [s], "aggregates": [{"kind": K, "input": s, "output_type": T}],
//    "children": [node]}
// T is a GpuValueType name ("INT64", ...) or its number; K a GpuAggregateKind name or number.
#pragma once

#include "json.h"

#include <memory>
#include <stdexcept>
#include <string>

namespace vgpu {
namespace plan_json {

inline int EnumValue(const json::Value &v, const char *const *names, int count, const char *what) {
	if (v.kind == json::Value::NUM) {
		int x = static_cast<int>(v.num);
		if (x < 0 || x >= count) {
			throw std::runtime_error(std::string("plan JSON: ") + what + " out of range");
		}
		return x;
	}
	if (v.kind == json::Value::STR) {
		for (int i = 0; i < count; i++) {
			if (v.str == names[i]) {
				return i;
			}
		}
		throw std::runtime_error(std::string("plan JSON: unknown ") + what + " '" + v.str + "'");
	}
	throw std::runtime_error(std::string("plan JSON: ") + what + " must be a name or a number");
}

inline vector_gpu::GpuValueType TypeOf(const json::Value &v) {
	static const char *const names[] = {"INT16",   "INT32",   "INT64",           "FLOAT32",
	                                    "FLOAT64", "BOOLEAN", "DICTIONARY_STRING", "HUGEINT"};
	return static_cast<vector_gpu::GpuValueType>(EnumValue(v, names, 8, "value type"));
}

inline vector_gpu::GpuOpType OpOf(const json::Value &v) {
	static const char *const names[] = {"SCAN", "FILTER", "PROJECTION", "HASH_JOIN", "GROUP_BY_AGGREGATE",
	                                    "CROSS_PRODUCT"};
	return static_cast<vector_gpu::GpuOpType>(EnumValue(v, names, 6, "op"));
}

inline vector_gpu::GpuAggregateKind AggOf(const json::Value &v) {
	static const char *const names[] = {"COUNT_STAR", "COUNT", "SUM", "MIN", "MAX", "AVG"};
	return static_cast<vector_gpu::GpuAggregateKind>(EnumValue(v, names, 6, "aggregate kind"));
}

inline std::shared_ptr<vector_gpu::GpuPlanNode> NodeFromJson(const json::Value &v, int depth = 0) {
	if (v.kind != json::Value::OBJ) {
		throw std::runtime_error("plan JSON: a plan node must be an object");
	}
	if (depth > 256) {
		throw std::runtime_error("plan JSON: plan nested too deeply");
	}
	if (!v.Has("op")) {
		throw std::runtime_error("plan JSON: node without \"op\"");
	}
	auto node = std::make_shared<vector_gpu::GpuPlanNode>();
	node->op_type = OpOf(v.At("op"));
	if (v.Has("mode")) {
		node->execution_mode = static_cast<vector_gpu::ExecutionMode>(static_cast<int>(v.Num("mode")));
	}
	node->estimated_rows = v.U64("estimated_rows", 0);
	node->estimated_row_width_bytes = static_cast<uint32_t>(v.U64("estimated_row_width_bytes", 0));
	const json::Value &table = v.At("table");
	node->table.catalog = table.Str("catalog");
	node->table.schema = table.Str("schema");
	node->table.table = table.Str("table");
	node->column_names = v.StrList("column_names");
	for (auto &e : v.At("expressions").arr) {
		vector_gpu::GpuExpr expr;
		expr.generated_cuda_source = e.Str("source");
		expr.source_hash = e.Str("hash");
		expr.input_columns = e.StrList("inputs");
		if (e.Has("output_type")) {
			expr.output_type = TypeOf(e.At("output_type"));
		}
		if (e.Str("window") == "RUNNING_SUM") {
			expr.window_primitive = vector_gpu::GpuWindowPrimitive::RUNNING_SUM;
		}
		expr.estimated_ops_per_row = e.Num("ops", 0.0);
		node->expressions.push_back(expr);
	}
	node->build_keys = v.StrList("build_keys");
	node->probe_keys = v.StrList("probe_keys");
	node->output_columns = v.StrList("output_columns");
	node->group_keys = v.StrList("group_keys");
	node->aggregate_exprs = v.StrList("aggregate_exprs");
	for (auto &a : v.At("aggregates").arr) {
		vector_gpu::GpuAggregate agg;
		agg.kind = AggOf(a.At("kind"));
		agg.input_column = a.Str("input");
		if (a.Has("output_type")) {
			agg.output_type = TypeOf(a.At("output_type"));
		}
		node->aggregates.push_back(agg);
	}
	for (auto &c : v.At("children").arr) {
		node->children.push_back(NodeFromJson(c, depth + 1));
	}
	return node;
}

inline std::shared_ptr<vector_gpu::GpuPlanNode> ParsePlan(const std::string &text) {
	return NodeFromJson(json::Parse(text));
}

} // namespace plan_json
} // namespace vgpu
