// gpu_filter_ops.cpp -- FilterGreaterThanInt32 (contract of cuda_engine/src/operators/gpu_filter.cu:36-85).
#include "../../../include/gpu_filter_ops.hpp"
#include "../host/device_pool.h"
#include "../hip/kernels.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace vector_gpu {

void FilterGreaterThanInt32(const int32_t *input, size_t n, int32_t threshold, int32_t *output, size_t *out_count) {
	if (out_count == nullptr) {
		throw std::invalid_argument("FilterGreaterThanInt32: out_count must not be null");
	}
	if (n == 0) {
		*out_count = 0;
		return;
	}
	if (input == nullptr || output == nullptr) {
		throw std::invalid_argument("FilterGreaterThanInt32: input/output must not be null when n > 0");
	}
	if (input == output) {
		throw std::invalid_argument("FilterGreaterThanInt32: in-place compaction (input == output) is not "
		                            "supported -- input and output must not overlap");
	}
	(void)hipGetLastError();
	try {
		vgpu::PoolArena arena(nullptr);
		auto *count = static_cast<uint64_t *>(arena.Alloc(sizeof(uint64_t)));
		vgpu::FilterGtI32(input, n, threshold, output, count, arena, nullptr);
		uint64_t host = 0;
		VGPU_HIP(hipMemcpy(&host, count, sizeof(host), hipMemcpyDeviceToHost));
		*out_count = static_cast<size_t>(host);
	} catch (const vgpu::HipError &e) {
		throw std::runtime_error(std::string("FilterGreaterThanInt32: ") + e.what());
	}
}

} // namespace vector_gpu
