// cost_model.cpp -- routing inputs declared in gpu_engine.hpp: row independence, plan shape and work
// density, working-set estimate, scanned-column size and the cached/uncached pipeline choice.
//
// IsRowIndependent, AnalyzePlanShape, EstimateColumnsSize and DetermineExecutionMode decide routing in
// the extension, so they follow cuda_engine/src/gpu_cost_model.cpp exactly (saturating u64 arithmetic,
// 8-byte default row width, FILTER ops counted over its input rows, density thresholds
// kMinOpsPerScannedRow / kOffloadOpsPerRow). EstimateWorkingSet models THIS engine's allocations, which
// differ from the reference's operator by operator:
//   SCAN          rows x width                     uploaded columns
//   FILTER        in_rows x width + tile state     fused kernel writes only the surviving rows of the
//                                                  carried columns (no bool masks, no u64 index vector)
//   PROJECTION    rows x width                     fused outputs
//   GROUP BY      in_rows x (16 + 16 per distinct  packed keys and payload, double-buffered by the radix
//                 aggregate input) + in_rows x     sort; outputs sized at the group bound (at most the
//                 (width + 8)                      input rows) plus the count state; a global aggregate
//                                                  needs one output row and no sort
//   HASH_JOIN     (left + right) x 48 +            per side: sorted key + row-id double buffers and its
//                 out_rows x (8 + width)           distinct-key runs; u32 pair indices and gathered outputs
//                 under a keyless GROUP BY:        aggregated without materializing: the smaller side sorted
//                 larger x 8 + smaller x 32        with its payload into a per-key table, the larger side
//                                                  regrouped by key range once (COUNT / integer SUM)
//                 (left + right) x 48              both sides sorted and merged (other aggregates)
//   CROSS_PRODUCT out_rows x width                 expanded outputs
#include "../../../include/gpu_engine.hpp"

#include <cstdint>
#include <limits>
#include <set>
#include <string>

namespace vector_gpu {

namespace {

constexpr uint64_t kMax = std::numeric_limits<uint64_t>::max();

uint64_t SatMul(uint64_t a, uint64_t b) {
	if (a == 0 || b == 0) {
		return 0;
	}
	return a > kMax / b ? kMax : a * b;
}

uint64_t SatAdd(uint64_t a, uint64_t b) {
	return a > kMax - b ? kMax : a + b;
}

uint64_t RowWidth(const GpuPlanNode &node) {
	return node.estimated_row_width_bytes == 0 ? 8 : node.estimated_row_width_bytes;
}

uint64_t NodeRows(const GpuPlanNode &node, uint64_t cap) {
	uint64_t rows = node.estimated_rows;
	return (cap != 0 && rows > cap) ? cap : rows;
}

uint64_t ChildRows(const GpuPlanNode &node, size_t index, uint64_t cap) {
	return index < node.children.size() && node.children[index] ? NodeRows(*node.children[index], cap) : 0;
}

//! How a join under a keyless GROUP BY runs: not aggregated (materialized), as the per-key table probed by
//! the larger side (every aggregate a COUNT or an integer SUM, whose output is HUGEINT), or sort-merge.
enum JoinAggregation { kMaterialized = 0, kProbeAggregate = 1, kSortAggregate = 2 };

JoinAggregation AggregationOf(const GpuPlanNode &node) {
	if (node.op_type != GpuOpType::GROUP_BY_AGGREGATE || !node.group_keys.empty()) {
		return kMaterialized;
	}
	for (auto &a : node.aggregates) {
		const bool count = a.kind == GpuAggregateKind::COUNT_STAR || a.kind == GpuAggregateKind::COUNT;
		const bool int_sum = a.kind == GpuAggregateKind::SUM && a.output_type == GpuValueType::HUGEINT;
		if (!count && !int_sum) {
			return kSortAggregate;
		}
	}
	return kProbeAggregate;
}

uint64_t NodeBytes(const GpuPlanNode &node, uint64_t cap, JoinAggregation aggregated) {
	const uint64_t width = RowWidth(node);
	const uint64_t out_rows = NodeRows(node, cap);
	switch (node.op_type) {
	case GpuOpType::FILTER: {
		const uint64_t in_rows = ChildRows(node, 0, cap);
		const uint64_t tiles = in_rows / 4096 + 1;
		return SatAdd(SatMul(in_rows, width), SatMul(tiles, 8));
	}
	case GpuOpType::GROUP_BY_AGGREGATE: {
		const uint64_t in_rows = ChildRows(node, 0, cap);
		std::set<std::string> inputs;
		for (auto &aggregate : node.aggregates) {
			if (aggregate.kind != GpuAggregateKind::COUNT_STAR && aggregate.kind != GpuAggregateKind::COUNT) {
				inputs.insert(aggregate.input_column);
			}
		}
		if (node.group_keys.empty()) {
			// Global aggregate: no sort, the reduction reads the inputs in place; one output row.
			return SatAdd(width, 8);
		}
		// Outputs are allocated at the group bound before the reduction runs; the bound is only known
		// from the keys' value ranges at run time, so the model takes its ceiling, the input row count
		// (estimated_rows, the expected group count, can be far below it).
		(void)out_rows;
		const uint64_t per_row = 16 + 16 * inputs.size();
		return SatAdd(SatMul(in_rows, per_row), SatMul(in_rows, SatAdd(width, 8)));
	}
	case GpuOpType::HASH_JOIN: {
		const uint64_t a = ChildRows(node, 0, cap), b = ChildRows(node, 1, cap);
		if (aggregated == kProbeAggregate) {
			// Under a keyless aggregate the join is never materialized: the smaller side is sorted with its
			// payload and laid out as a per-key table, the larger side is read once (grouped by key range
			// into 4-byte keys + payload).
			const uint64_t small = a < b ? a : b, large = a < b ? b : a;
			return SatAdd(SatMul(large, 8), SatMul(small, 32));
		}
		if (aggregated == kSortAggregate) {
			// The other aggregates (MIN/MAX, floating SUM, AVG) sort both sides with their payload and merge
			// the runs; nothing is materialized.
			return SatMul(SatAdd(a, b), 48);
		}
		return SatAdd(SatMul(SatAdd(a, b), 48), SatMul(out_rows, SatAdd(8, width)));
	}
	case GpuOpType::SCAN:
	case GpuOpType::PROJECTION:
	case GpuOpType::CROSS_PRODUCT:
	default:
		return SatMul(out_rows, width);
	}
}

const char *OperatorName(GpuOpType op) {
	switch (op) {
	case GpuOpType::SCAN:
		return "SCAN";
	case GpuOpType::FILTER:
		return "FILTER";
	case GpuOpType::PROJECTION:
		return "PROJECTION";
	case GpuOpType::HASH_JOIN:
		return "HASH_JOIN";
	case GpuOpType::GROUP_BY_AGGREGATE:
		return "GROUP_BY_AGGREGATE";
	case GpuOpType::CROSS_PRODUCT:
		return "CROSS_PRODUCT";
	}
	return "UNKNOWN";
}

//! `aggregated`: how the node's parent, when it is a GROUP BY without keys, aggregates a join below it.
void Accumulate(const GpuPlanNode &node, uint64_t cap, JoinAggregation aggregated, GpuWorkingSetEstimate &out) {
	const JoinAggregation below = AggregationOf(node);
	for (auto &child : node.children) {
		if (child) {
			Accumulate(*child, cap, below, out);
		}
	}
	const uint64_t bytes = NodeBytes(node, cap, aggregated);
	out.total_bytes = SatAdd(out.total_bytes, bytes);
	if (bytes > out.largest_bytes) {
		out.largest_bytes = bytes;
		out.largest_operator = OperatorName(node.op_type);
	}
}

double NodeOps(const GpuPlanNode &node) {
	double per_row = 0.0;
	for (auto &expr : node.expressions) {
		per_row += expr.estimated_ops_per_row;
	}
	if (per_row == 0.0) {
		return 0.0;
	}
	// A FILTER evaluates its predicates on every input row, not only on the rows it keeps.
	const uint64_t rows = node.op_type == GpuOpType::FILTER ? ChildRows(node, 0, 0) : node.estimated_rows;
	return per_row * static_cast<double>(rows);
}

void AccumulateShape(const GpuPlanNode &node, GpuPlanShape &shape) {
	for (auto &child : node.children) {
		if (child) {
			AccumulateShape(*child, shape);
		}
	}
	if (node.op_type == GpuOpType::SCAN) {
		shape.scanned_rows = SatAdd(shape.scanned_rows, node.estimated_rows);
		shape.scanned_bytes = SatAdd(shape.scanned_bytes, SatMul(node.estimated_rows, RowWidth(node)));
	}
	if (node.estimated_rows > shape.max_rows) {
		shape.max_rows = node.estimated_rows;
	}
	shape.total_ops += NodeOps(node);
}

GpuWorkDensity ClassifyDensity(const GpuPlanShape &shape) {
	if (shape.total_ops == 0.0) {
		return GpuWorkDensity::INCONCLUSIVE;
	}
	if (shape.ops_per_scanned_row >= kOffloadOpsPerRow) {
		return GpuWorkDensity::SUFFICIENT;
	}
	if (shape.ops_per_scanned_row < kMinOpsPerScannedRow) {
		return GpuWorkDensity::INSUFFICIENT;
	}
	return GpuWorkDensity::INCONCLUSIVE;
}

} // namespace

bool IsRowIndependent(const GpuPlanNode &node) {
	if (node.op_type == GpuOpType::GROUP_BY_AGGREGATE || node.op_type == GpuOpType::HASH_JOIN ||
	    node.op_type == GpuOpType::CROSS_PRODUCT) {
		return false;
	}
	for (auto &child : node.children) {
		if (child && !IsRowIndependent(*child)) {
			return false;
		}
	}
	return true;
}

GpuPlanShape AnalyzePlanShape(const GpuPlanNode &plan) {
	GpuPlanShape shape;
	AccumulateShape(plan, shape);
	shape.amplification = shape.scanned_rows == 0 ? static_cast<double>(shape.max_rows)
	                                               : static_cast<double>(shape.max_rows) /
	                                                     static_cast<double>(shape.scanned_rows);
	shape.ops_per_scanned_row =
	    shape.scanned_rows == 0 ? shape.total_ops : shape.total_ops / static_cast<double>(shape.scanned_rows);
	shape.ops_per_scanned_byte =
	    shape.scanned_bytes == 0 ? shape.total_ops : shape.total_ops / static_cast<double>(shape.scanned_bytes);
	shape.density = ClassifyDensity(shape);
	return shape;
}

GpuWorkingSetEstimate EstimateWorkingSet(const GpuPlanNode &plan, uint64_t max_chunk_rows) {
	GpuWorkingSetEstimate estimate;
	const bool chunked = max_chunk_rows != 0 && IsRowIndependent(plan);
	estimate.chunk_capped = chunked;
	Accumulate(plan, chunked ? max_chunk_rows : 0, kMaterialized, estimate);
	return estimate;
}

uint64_t EstimateColumnsSize(const GpuPlanNode &plan) {
	uint64_t total = 0;
	if (plan.op_type == GpuOpType::SCAN) {
		total = SatMul(plan.estimated_rows, RowWidth(plan));
	}
	for (const auto &child : plan.children) {
		if (child) {
			total = SatAdd(total, EstimateColumnsSize(*child));
		}
	}
	return total;
}

ExecutionMode DetermineExecutionMode(const GpuPlanNode &plan, uint64_t vram_budget, double threshold) {
	const uint64_t columns = EstimateColumnsSize(plan);
	if (vram_budget > 0 && static_cast<double>(columns) > static_cast<double>(vram_budget) * threshold) {
		return ExecutionMode::PIPELINE_STREAM_UNCACHED;
	}
	return ExecutionMode::PIPELINE_STREAM_CACHED;
}

} // namespace vector_gpu
