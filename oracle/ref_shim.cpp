// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY. extern "C" entry points over the reference's own,
// device-independent sources (cuda_engine/src/gpu_cost_model.cpp and cuda_engine/src/kernel_cache.cpp,
// compiled unmodified from the reference tree by oracle/build_ref.py into oracle/_ref/libvref.so), so
// tests can compare this engine's routing functions and SHA-256 with the reference's on the same plans.
//
// Compiled against the REFERENCE's gpu_engine.hpp (include path set by build_ref.py); plan_json.h only
// uses the names that header declares, so plans are built into the reference's own structs.
#include "gpu_engine.hpp"
#include "kernel_cache.hpp"

#include "../vectisdb_amd/csrc/engine/plan_json.h"

#include <cstring>
#include <exception>
#include <string>

using namespace vector_gpu;

extern "C" {

//! out: scanned_rows, scanned_bytes, max_rows, total_ops, ops_per_scanned_row, ops_per_scanned_byte,
//! density, amplification (as doubles). Returns 0, or 1 on a parse error.
int vref_analyze_plan_shape(const char *json, double *out) {
	try {
		auto s = AnalyzePlanShape(*vgpu::plan_json::ParsePlan(json));
		out[0] = (double)s.scanned_rows;
		out[1] = (double)s.scanned_bytes;
		out[2] = (double)s.max_rows;
		out[3] = s.total_ops;
		out[4] = s.ops_per_scanned_row;
		out[5] = s.ops_per_scanned_byte;
		out[6] = (double)(int)s.density;
		out[7] = s.amplification;
		return 0;
	} catch (const std::exception &) {
		return 1;
	}
}

//! Exact integer fields of the shape (the doubles above lose precision past 2^53).
int vref_plan_shape_counts(const char *json, uint64_t *out) {
	try {
		auto s = AnalyzePlanShape(*vgpu::plan_json::ParsePlan(json));
		out[0] = s.scanned_rows;
		out[1] = s.scanned_bytes;
		out[2] = s.max_rows;
		return 0;
	} catch (const std::exception &) {
		return 1;
	}
}

int vref_is_row_independent(const char *json) {
	try {
		return IsRowIndependent(*vgpu::plan_json::ParsePlan(json)) ? 1 : 0;
	} catch (const std::exception &) {
		return -1;
	}
}

int vref_estimate_columns_size(const char *json, uint64_t *out) {
	try {
		*out = EstimateColumnsSize(*vgpu::plan_json::ParsePlan(json));
		return 0;
	} catch (const std::exception &) {
		return 1;
	}
}

int vref_determine_execution_mode(const char *json, uint64_t budget, double threshold) {
	try {
		return (int)DetermineExecutionMode(*vgpu::plan_json::ParsePlan(json), budget, threshold);
	} catch (const std::exception &) {
		return -1;
	}
}

int vref_estimate_working_set(const char *json, uint64_t max_chunk_rows, uint64_t *total, int *capped, char *name,
                              size_t cap) {
	try {
		auto e = EstimateWorkingSet(*vgpu::plan_json::ParsePlan(json), max_chunk_rows);
		*total = e.total_bytes;
		*capped = e.chunk_capped ? 1 : 0;
		std::strncpy(name, e.largest_operator, cap - 1);
		name[cap - 1] = '\0';
		return 0;
	} catch (const std::exception &) {
		return 1;
	}
}

void vref_hash_source(const char *text, size_t len, char *hex) {
	std::string h = KernelCache::HashSource(std::string(text, len));
	std::memcpy(hex, h.c_str(), h.size() + 1);
}

} // extern "C"
