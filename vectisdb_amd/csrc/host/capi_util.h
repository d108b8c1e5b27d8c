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
