// plan_capi.cpp -- extern "C" entry points for whole plans (GpuEngine::ExecutePlan), the routing cost
// model, the column cache, the device budget and the kernel-cache hash.
#include "../../../include/code_generator.hpp"
#include "../../../include/gpu_column_cache.hpp"
#include "../../../include/gpu_engine.hpp"
#include "../../../include/gpu_semantics.hpp"
#include "../../../include/kernel_cache.hpp"
#include "../../../include/rmm_pool.hpp"
#include "../host/capi_util.h"
#include "plan_json.h"

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct vgpu_result {
	vector_gpu::GpuExecutionResult result;
};

using namespace vgpu;

namespace {

vector_gpu::GpuTableRef TableRef(const char *catalog, const char *schema, const char *table) {
	if (catalog == nullptr || schema == nullptr || table == nullptr) {
		throw std::runtime_error("table reference parts must not be NULL");
	}
	return vector_gpu::GpuTableRef {catalog, schema, table};
}

std::shared_ptr<vector_gpu::GpuPlanNode> Plan(const char *text) {
	if (text == nullptr) {
		throw std::runtime_error("plan JSON is NULL");
	}
	return plan_json::ParsePlan(text);
}

} // namespace

extern "C" {

int vgpu_execute_plan(const char *plan_json, const vgpu_host_column *inputs, int32_t ninputs,
                      const vgpu_execute_options *options, vgpu_result **result) {
	return Guard([&] {
		if (result == nullptr) {
			throw std::runtime_error("vgpu_execute_plan: result must not be NULL");
		}
		*result = nullptr;
		auto plan = Plan(plan_json);
		std::vector<vector_gpu::GpuColumn> columns;
		for (int32_t i = 0; i < ninputs; i++) {
			vector_gpu::GpuColumn c;
			c.name = inputs[i].name != nullptr ? inputs[i].name : "";
			c.type = static_cast<vector_gpu::GpuValueType>(inputs[i].type);
			c.data = inputs[i].data;
			c.validity = inputs[i].validity;
			c.row_count = inputs[i].rows;
			columns.push_back(std::move(c));
		}
		vector_gpu::GpuExecuteOptions opts;
		if (options != nullptr) {
			opts.stage_scan_to_cache = options->stage_scan_to_cache != 0;
			opts.cache_chunk_index = options->cache_chunk_index;
			opts.serve_whole_scan_from_cache = options->serve_whole_scan_from_cache != 0;
		}
		auto *r = new vgpu_result();
		r->result = vector_gpu::GpuEngine::Instance().ExecutePlan(*plan, columns, opts);
		*result = r;
	});
}

int32_t vgpu_result_success(const vgpu_result *result) {
	return result != nullptr && result->result.success ? 1 : 0;
}

const char *vgpu_result_error(const vgpu_result *result) {
	return result == nullptr ? "no result" : result->result.error_message.c_str();
}

int32_t vgpu_result_num_columns(const vgpu_result *result) {
	return result == nullptr ? 0 : static_cast<int32_t>(result->result.columns.size());
}

int vgpu_result_column(const vgpu_result *result, int32_t index, vgpu_host_column *out) {
	return Guard([&] {
		if (result == nullptr || out == nullptr || index < 0 ||
		    index >= static_cast<int32_t>(result->result.columns.size())) {
			throw std::runtime_error("vgpu_result_column: bad result or column index");
		}
		const auto &c = result->result.columns[static_cast<size_t>(index)];
		out->name = c.name.c_str();
		out->type = static_cast<int32_t>(c.type);
		out->data = c.data;
		out->validity = c.validity;
		out->rows = c.row_count;
	});
}

void vgpu_result_free(vgpu_result *result) {
	delete result;
}

int vgpu_analyze_plan_shape(const char *plan_json, vgpu_plan_shape *out) {
	return Guard([&] {
		auto shape = vector_gpu::AnalyzePlanShape(*Plan(plan_json));
		out->scanned_rows = shape.scanned_rows;
		out->scanned_bytes = shape.scanned_bytes;
		out->max_rows = shape.max_rows;
		out->total_ops = shape.total_ops;
		out->ops_per_scanned_row = shape.ops_per_scanned_row;
		out->ops_per_scanned_byte = shape.ops_per_scanned_byte;
		out->density = static_cast<int32_t>(shape.density);
		out->amplification = shape.amplification;
	});
}

int vgpu_is_row_independent(const char *plan_json, int32_t *out) {
	return Guard([&] { *out = vector_gpu::IsRowIndependent(*Plan(plan_json)) ? 1 : 0; });
}

int vgpu_estimate_working_set(const char *plan_json, uint64_t max_chunk_rows, uint64_t *total_bytes,
                              int32_t *chunk_capped, uint64_t *largest_bytes, char *largest_operator, size_t cap) {
	return Guard([&] {
		auto e = vector_gpu::EstimateWorkingSet(*Plan(plan_json), max_chunk_rows);
		*total_bytes = e.total_bytes;
		*chunk_capped = e.chunk_capped ? 1 : 0;
		*largest_bytes = e.largest_bytes;
		if (largest_operator != nullptr && cap > 0) {
			std::strncpy(largest_operator, e.largest_operator, cap - 1);
			largest_operator[cap - 1] = '\0';
		}
	});
}

int vgpu_estimate_columns_size(const char *plan_json, uint64_t *out) {
	return Guard([&] { *out = vector_gpu::EstimateColumnsSize(*Plan(plan_json)); });
}

int vgpu_determine_execution_mode(const char *plan_json, uint64_t vram_budget, double threshold, int32_t *mode) {
	return Guard([&] {
		*mode = static_cast<int32_t>(vector_gpu::DetermineExecutionMode(*Plan(plan_json), vram_budget, threshold));
	});
}

int vgpu_hash_source(const char *text, size_t len, char *hex) {
	return Guard([&] {
		std::string digest = vector_gpu::KernelCache::HashSource(std::string(text == nullptr ? "" : text, len));
		std::memcpy(hex, digest.c_str(), digest.size() + 1);
	});
}

int32_t vgpu_fast_math_enabled(void) {
	return vector_gpu::IsGpuFastMathEnabled() ? 1 : 0;
}

void vgpu_set_fast_math(int32_t on) {
	vector_gpu::SetGpuFastMathEnabled(on != 0);
}

int32_t vgpu_extended_semantics_enabled(void) {
	return vector_gpu::IsGpuExtendedSemanticsEnabled() ? 1 : 0;
}

void vgpu_set_extended_semantics(int32_t on) {
	vector_gpu::SetGpuExtendedSemanticsEnabled(on != 0);
}

int vgpu_cache_begin_staging(const char *catalog, const char *schema, const char *table) {
	return Guard([&] { vector_gpu::GpuColumnCache::Instance().BeginStaging(TableRef(catalog, schema, table)); });
}

int vgpu_cache_commit_staging(const char *catalog, const char *schema, const char *table, uint64_t scanned_rows) {
	return Guard([&] {
		vector_gpu::GpuColumnCache::Instance().CommitStaging(TableRef(catalog, schema, table), scanned_rows);
	});
}

int vgpu_cache_abort_staging(const char *catalog, const char *schema, const char *table) {
	return Guard([&] { vector_gpu::GpuColumnCache::Instance().AbortStaging(TableRef(catalog, schema, table)); });
}

int vgpu_cache_invalidate_table(const char *catalog, const char *schema, const char *table) {
	return Guard([&] { vector_gpu::GpuColumnCache::Instance().InvalidateTable(TableRef(catalog, schema, table)); });
}

int vgpu_cache_invalidate_all(void) {
	return Guard([&] { vector_gpu::GpuColumnCache::Instance().InvalidateAll(); });
}

int vgpu_cache_current_bytes(uint64_t *out) {
	return Guard([&] { *out = vector_gpu::GpuColumnCache::Instance().CurrentBytes(); });
}

int vgpu_cache_table_generation(const char *catalog, const char *schema, const char *table, uint64_t *out) {
	return Guard(
	    [&] { *out = vector_gpu::GpuColumnCache::Instance().TableGeneration(TableRef(catalog, schema, table)); });
}

int vgpu_cache_replayable_chunk_rows(const char *catalog, const char *schema, const char *table,
                                     const char *const *columns, int32_t ncolumns, uint64_t *rows, size_t cap,
                                     size_t *nchunks) {
	return Guard([&] {
		std::vector<std::string> names;
		for (int32_t i = 0; i < ncolumns; i++) {
			names.emplace_back(columns[i]);
		}
		auto chunk_rows =
		    vector_gpu::GpuColumnCache::Instance().ReplayableChunkRows(TableRef(catalog, schema, table), names);
		*nchunks = chunk_rows.size();
		if (chunk_rows.size() > cap) {
			throw std::runtime_error("vgpu_cache_replayable_chunk_rows: more chunks than `cap`");
		}
		for (size_t i = 0; i < chunk_rows.size(); i++) {
			rows[i] = chunk_rows[i];
		}
	});
}

int vgpu_vram_budget_init(double fraction_of_vram) {
	return Guard([&] { vector_gpu::RmmPool::Instance().EnsureInitialized(fraction_of_vram); });
}

int vgpu_vram_has_headroom(uint64_t bytes, int32_t *out) {
	return Guard([&] { *out = vector_gpu::RmmPool::Instance().HasHeadroomFor(bytes) ? 1 : 0; });
}

int vgpu_vram_budget_bytes(uint64_t *total, uint64_t *allocated) {
	return Guard([&] {
		*total = vector_gpu::RmmPool::Instance().TotalPoolBytes();
		*allocated = vector_gpu::RmmPool::Instance().AllocatedBytes();
	});
}

} // extern "C"
