#pragma once
// gpu_engine.hpp -- the engine API the DuckDB extension calls (drop-in for cuda_engine/include/gpu_engine.hpp).
//
// Same namespace, type names, fields, enum values and function signatures as the reference header, so
// extension/ recompiles unchanged against it. No HIP or DuckDB type appears here: the extension converts
// DuckDB vectors to GpuColumn and plans to GpuPlanNode; everything device-side lives behind ExecutePlan.

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace vector_gpu {

enum class GpuOpType : uint8_t {
	SCAN,
	FILTER,
	PROJECTION,
	HASH_JOIN,
	GROUP_BY_AGGREGATE,
	CROSS_PRODUCT,
};

enum class ExecutionMode : uint8_t {
	PIPELINE_STREAM_CACHED = 0,
	PIPELINE_STREAM_UNCACHED = 1,
	BATCH_DISPATCH = 2,
};

enum class GpuValueType : uint8_t {
	INT16,
	INT32,
	INT64,
	FLOAT32,
	FLOAT64,
	BOOLEAN,
	DICTIONARY_STRING, // index into GpuColumn::string_dictionary; not executable on the device
	HUGEINT,           // 128-bit two's complement, laid out as GpuHugeInt
};

//! DuckDB's hugeint_t layout: low 64 bits unsigned, high 64 bits signed.
struct alignas(16) GpuHugeInt {
	uint64_t lower;
	int64_t upper;

	bool operator==(const GpuHugeInt &other) const {
		return lower == other.lower && upper == other.upper;
	}
};

enum class GpuWindowPrimitive {
	NONE,
	RUNNING_SUM,
};

//! One scalar expression as source text. The body is a statement that reads
//! `((const T*)inputs[k])[row]` and writes `((U*)out)[idx]`; the engine compiles it with hipRTC.
struct GpuExpr {
	std::string source_hash;
	std::string generated_cuda_source;
	//! Column feeding inputs[k], by name, in slot order.
	std::vector<std::string> input_columns;
	GpuValueType output_type = GpuValueType::FLOAT64;
	GpuWindowPrimitive window_primitive = GpuWindowPrimitive::NONE;
	//! Weighted ops per row measured by the extension's cost visitor; 0 = unmeasured.
	double estimated_ops_per_row = 0.0;
};

enum class GpuAggregateKind : uint8_t {
	COUNT_STAR,
	COUNT,
	SUM,
	MIN,
	MAX,
	AVG,
};

struct GpuAggregate {
	GpuAggregateKind kind = GpuAggregateKind::COUNT_STAR;
	//! Input column name as produced by the child; empty for COUNT_STAR.
	std::string input_column;
	//! COUNT -> INT64, MIN/MAX -> input type, SUM -> FLOAT64 (float input) or HUGEINT (integer input),
	//! AVG -> FLOAT64.
	GpuValueType output_type = GpuValueType::INT64;
};

//! Fully qualified table identity (catalog, schema, table).
struct GpuTableRef {
	std::string catalog;
	std::string schema;
	std::string table;

	bool operator==(const GpuTableRef &other) const {
		return catalog == other.catalog && schema == other.schema && table == other.table;
	}

	//! For logs and messages only (never a map key: identifiers may contain dots).
	std::string ToString() const {
		return catalog + "." + schema + "." + table;
	}
};

struct GpuPlanNode {
	GpuOpType op_type = GpuOpType::SCAN;
	ExecutionMode execution_mode = ExecutionMode::PIPELINE_STREAM_CACHED;
	std::vector<std::shared_ptr<GpuPlanNode>> children;

	// Cost-model inputs (advisory). For GROUP_BY_AGGREGATE, estimated_rows is the group count.
	uint64_t estimated_rows = 0;
	uint32_t estimated_row_width_bytes = 0;

	// SCAN
	GpuTableRef table;
	std::vector<std::string> column_names;

	// FILTER / PROJECTION
	std::vector<GpuExpr> expressions;

	// HASH_JOIN: build_keys name children[0] columns, probe_keys children[1] columns.
	std::vector<std::string> build_keys;
	std::vector<std::string> probe_keys;
	//! Exact output column list, in order (HASH_JOIN and CROSS_PRODUCT).
	std::vector<std::string> output_columns;

	// GROUP_BY_AGGREGATE: output = group keys, then aggregates named agg_<i>.
	std::vector<std::string> group_keys;
	std::vector<std::string> aggregate_exprs;
	std::vector<GpuAggregate> aggregates;
};

//! Host column handed across the boundary: pageable memory owned by the caller, tightly packed.
//! validity: Arrow bitmap (bit r%8 of byte r/8, 1 = valid) or nullptr when the column has no NULLs.
struct GpuColumn {
	std::string name;
	GpuValueType type;
	const void *data = nullptr;
	const uint8_t *validity = nullptr;
	uint64_t row_count = 0;
	std::vector<std::string> string_dictionary;
};

struct GpuExecutionResult {
	bool success = false;
	std::string error_message;
	std::vector<GpuColumn> columns;
	//! Backing storage of `columns` (host memory owned by the result).
	std::vector<std::vector<uint8_t>> owned_buffers;
};

//! Predicted peak device bytes of one plan execution (saturating).
struct GpuWorkingSetEstimate {
	uint64_t total_bytes = 0;
	bool chunk_capped = false;
	const char *largest_operator = "";
	uint64_t largest_bytes = 0;
};

//! True when each output row depends only on its own input row (SCAN/FILTER/PROJECTION trees).
bool IsRowIndependent(const GpuPlanNode &node);

enum class GpuWorkDensity : uint8_t {
	INSUFFICIENT,
	INCONCLUSIVE,
	SUFFICIENT,
};

//! Routing thresholds on weighted ops per scanned row.
constexpr double kMinOpsPerScannedRow = 3.0;
constexpr double kOffloadOpsPerRow = 10.0;

struct GpuPlanShape {
	uint64_t scanned_rows = 0;
	uint64_t scanned_bytes = 0;
	uint64_t max_rows = 0;
	double total_ops = 0.0;
	double ops_per_scanned_row = 0.0;
	double ops_per_scanned_byte = 0.0;
	GpuWorkDensity density = GpuWorkDensity::INSUFFICIENT;
	double amplification = 0.0;
};

GpuPlanShape AnalyzePlanShape(const GpuPlanNode &plan);

//! Peak device memory of `plan` on this engine (derived from the operators' own allocations). Row counts
//! of a row-independent plan are capped at `max_chunk_rows` (0 = no cap).
GpuWorkingSetEstimate EstimateWorkingSet(const GpuPlanNode &plan, uint64_t max_chunk_rows);

uint64_t EstimateColumnsSize(const GpuPlanNode &plan);

ExecutionMode DetermineExecutionMode(const GpuPlanNode &plan, uint64_t vram_budget, double threshold = 0.75);

//! Column-cache interaction of one execution. The two read modes and the staging mode are mutually
//! exclusive.
struct GpuExecuteOptions {
	bool stage_scan_to_cache = false;
	int64_t cache_chunk_index = -1;
	bool serve_whole_scan_from_cache = false;
};

class GpuEngine {
public:
	static GpuEngine &Instance();

	//! Executes `plan` over `inputs` (SCAN leaves consume inputs in traversal order) and returns host
	//! columns. Never throws: failures come back as success=false with a message.
	GpuExecutionResult ExecutePlan(const GpuPlanNode &plan, const std::vector<GpuColumn> &inputs,
	                               const GpuExecuteOptions &options = {});

private:
	GpuEngine() = default;
};

} // namespace vector_gpu
