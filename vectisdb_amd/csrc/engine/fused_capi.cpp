// fused_capi.cpp -- extern "C" entry points of the fused filter/projection and cross-product kernels.
#include "../host/capi_util.h"
#include "../host/device_pool.h"
#include "fused.h"

#include <stdexcept>
#include <string>
#include <vector>

using namespace vgpu;

namespace {

FusedExpr ToFused(const vgpu_expr &e, int32_t ncols) {
	FusedExpr f;
	if (e.body == nullptr) {
		throw std::runtime_error("expression body is NULL");
	}
	f.body = e.body;
	for (int32_t k = 0; k < e.ninputs; k++) {
		if (e.inputs[k] < 0 || e.inputs[k] >= ncols) {
			throw std::runtime_error("expression input index out of range");
		}
		f.slots.push_back(e.inputs[k]);
	}
	f.out_type = e.output_type;
	return f;
}

} // namespace

extern "C" {

int vgpu_filter_project(const vgpu_column *cols, const uint8_t *const *col_valid, int32_t ncols,
                        const vgpu_expr *preds, int32_t npreds, const vgpu_expr *projs, int32_t nprojs,
                        void *const *outs, uint8_t *const *out_valid, uint32_t *sel_out, uint64_t *kept,
                        void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		if (ncols < 0 || ncols > kFusedMaxCols || nprojs > kFusedMaxOuts) {
			throw std::runtime_error("vgpu_filter_project: at most 16 columns and 16 projections");
		}
		uint64_t rows = ncols > 0 ? cols[0].rows : 0;
		FusedSpec spec;
		std::vector<const void *> ptrs;
		std::vector<const uint8_t *> valids;
		for (int32_t c = 0; c < ncols; c++) {
			if (cols[c].rows != rows) {
				throw std::runtime_error("vgpu_filter_project: input column row count mismatch");
			}
			FusedColumn fc;
			fc.vtype = cols[c].type;
			fc.has_valid = col_valid != nullptr && col_valid[c] != nullptr;
			spec.cols.push_back(fc);
			ptrs.push_back(cols[c].data);
			valids.push_back(col_valid != nullptr ? col_valid[c] : nullptr);
		}
		for (int32_t i = 0; i < npreds; i++) {
			spec.predicates.push_back(ToFused(preds[i], ncols));
		}
		for (int32_t i = 0; i < nprojs; i++) {
			spec.projections.push_back(ToFused(projs[i], ncols));
		}
		spec.emit_selection = sel_out != nullptr;
		auto kernel = GetFusedKernel(spec);
		PoolArena arena(s);
		FusedLaunch la;
		la.cols = ptrs.data();
		la.valid = valids.data();
		la.outs = outs;
		la.out_valid = out_valid;
		la.n = rows;
		la.sel_out = sel_out;
		void *state = nullptr;
		size_t state_bytes = 0;
		if (npreds > 0) {
			la.count_device = static_cast<uint64_t *>(arena.Alloc(sizeof(uint64_t)));
			state_bytes = FusedKernel::StateBytes(rows);
			state = arena.Alloc(state_bytes);
		}
		kernel->Launch(la, state, state_bytes, s);
		if (npreds > 0) {
			VGPU_HIP(hipMemcpyAsync(kept, la.count_device, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
			if (rows == 0) {
				*kept = 0;
			}
		} else {
			*kept = rows;
		}
		VGPU_HIP(hipStreamSynchronize(s));
	});
}

int vgpu_cross_filter(const vgpu_column *cols, int32_t ncols, int32_t nleft, uint64_t left_rows, uint64_t right_rows,
                      const vgpu_expr *preds, int32_t npreds, const vgpu_expr *projs, int32_t nprojs,
                      void *const *outs, uint64_t capacity, uint64_t *kept, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		if (ncols < 0 || ncols > kFusedMaxCols || nprojs > kFusedMaxOuts || nleft < 0 || nleft > ncols) {
			throw std::runtime_error("vgpu_cross_filter: bad column counts");
		}
		constexpr uint64_t kMaxPairs = 1ull << 40;
		if (right_rows != 0 && left_rows > kMaxPairs / right_rows) {
			throw std::runtime_error("vgpu_cross_filter: cross product too large");
		}
		const uint64_t total = left_rows * right_rows;
		FusedSpec spec;
		spec.cross = true;
		std::vector<const void *> ptrs;
		for (int32_t c = 0; c < ncols; c++) {
			uint64_t expect = c < nleft ? left_rows : right_rows;
			if (cols[c].rows != expect) {
				throw std::runtime_error("vgpu_cross_filter: column row count does not match its side");
			}
			FusedColumn fc;
			fc.vtype = cols[c].type;
			fc.side = c < nleft ? 0 : 1;
			spec.cols.push_back(fc);
			ptrs.push_back(cols[c].data);
		}
		for (int32_t i = 0; i < npreds; i++) {
			spec.predicates.push_back(ToFused(preds[i], ncols));
		}
		for (int32_t i = 0; i < nprojs; i++) {
			spec.projections.push_back(ToFused(projs[i], ncols));
		}
		if (npreds == 0 && capacity < total) {
			throw std::runtime_error("vgpu_cross_filter: capacity below left_rows * right_rows");
		}
		auto kernel = GetFusedKernel(spec);
		PoolArena arena(s);
		FusedLaunch la;
		la.cols = ptrs.data();
		la.outs = outs;
		la.n = total;
		la.right_rows = right_rows;
		la.capacity = capacity;
		void *state = nullptr;
		size_t state_bytes = 0;
		if (npreds > 0) {
			la.count_device = static_cast<uint64_t *>(arena.Alloc(sizeof(uint64_t)));
			state_bytes = FusedKernel::StateBytes(total);
			state = arena.Alloc(state_bytes);
		}
		kernel->Launch(la, state, state_bytes, s);
		uint64_t count = total;
		if (npreds > 0 && total > 0) {
			VGPU_HIP(hipMemcpyAsync(&count, la.count_device, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
		} else if (npreds > 0) {
			count = 0;
		}
		VGPU_HIP(hipStreamSynchronize(s));
		*kept = count;
		if (count > capacity) {
			throw std::runtime_error("vgpu_cross_filter: " + std::to_string(count) +
			                         " pairs pass the filter, more than the output capacity " +
			                         std::to_string(capacity));
		}
	});
}

} // extern "C"
