// capi_util.h -- shared helpers of the extern "C" translation units.
#pragma once

#include "../../../include/vgpu.h"
#include "operators.h"

#include <exception>
#include <string>
#include <vector>

namespace vgpu {

void SetLastError(const std::string &message);
std::vector<DevCol> ToDevCols(const vgpu_column *cols, int32_t n);
//! The batched form's arguments from the C ABI's (offsets: device row offsets, or nullptr for callers that
//! pass the row groups another way).
BatchedGroupByArgs MakeBatchedArgs(const vgpu_column *cols, int32_t ncols, uint64_t rows, const uint64_t *offsets,
                                   int32_t nsegments, const vgpu_expr *preds, int32_t npreds, int32_t key_col,
                                   const int32_t *value_cols, int32_t nvalues, const vgpu_agg *aggs, int32_t naggs,
                                   void *segment_out, void *key_out, void *const *agg_out, uint64_t capacity,
                                   uint64_t *segment_group_offsets);

//! Runs `body`, mapping any exception to VGPU_ERROR + message; never lets one escape into C callers.
template <typename F>
int Guard(F &&body) {
	try {
		SetLastError("");
		body();
		return VGPU_OK;
	} catch (const std::exception &e) {
		SetLastError(e.what());
	} catch (...) {
		SetLastError("unknown C++ exception");
	}
	return VGPU_ERROR;
}

} // namespace vgpu
