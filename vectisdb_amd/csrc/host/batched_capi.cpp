// batched_capi.cpp -- extern "C" entry points of the batched form over row ranges: the eager pipeline
// (planned per batch on the host) and the prepared plan replayed as one HIP graph per batch.
#include "batched_graph.h"
#include "capi_util.h"
#include "device_pool.h"

#include <stdexcept>
#include <string>

struct vgpu_batched_plan {
	std::unique_ptr<vgpu::BatchedGraphPlan> plan;
};

using namespace vgpu;

extern "C" {

int vgpu_batched_ranges_filter_group_by(const vgpu_column *cols, int32_t ncols, uint64_t rows,
                                        const uint64_t *row_starts, const uint64_t *row_lens, int32_t nranges,
                                        const vgpu_expr *preds, int32_t npreds, int32_t key_col,
                                        const int32_t *value_cols, int32_t nvalues, const vgpu_agg *aggs,
                                        int32_t naggs, void *segment_out, void *key_out, void *const *agg_out,
                                        uint64_t capacity, uint64_t *groups, uint64_t *segment_group_offsets,
                                        void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		if (row_starts == nullptr || row_lens == nullptr) {
			throw std::runtime_error("batched group by: row starts and lengths are required");
		}
		auto a = MakeBatchedArgs(cols, ncols, rows, nullptr, nranges, preds, npreds, key_col, value_cols, nvalues,
		                         aggs, naggs, segment_out, key_out, agg_out, capacity, segment_group_offsets);
		a.range_rows_host = row_starts;
		a.range_lens_host = row_lens;
		*groups = BatchedFilterGroupByDevice(a, arena, s);
	});
}

int vgpu_batched_plan_create(const vgpu_column *cols, int32_t ncols, uint64_t rows, int32_t max_ranges,
                             uint64_t max_rows, const vgpu_expr *preds, int32_t npreds, int32_t key_col,
                             const int32_t *value_cols, int32_t nvalues, const vgpu_agg *aggs, int32_t naggs,
                             void *segment_out, void *key_out, void *const *agg_out, uint64_t capacity,
                             vgpu_batched_plan **plan) {
	return Guard([&] {
		if (plan == nullptr) {
			throw std::runtime_error("batched plan: no handle to write");
		}
		*plan = nullptr;
		auto a = MakeBatchedArgs(cols, ncols, rows, nullptr, max_ranges, preds, npreds, key_col, value_cols, nvalues,
		                         aggs, naggs, segment_out, key_out, agg_out, capacity, nullptr);
		BatchedPlanArgs p;
		p.cols = a.cols;
		p.table_rows = rows;
		p.max_queries = a.nseg;
		p.max_rows = max_rows;
		p.preds = a.preds;
		p.key_col = a.key_col;
		p.value_cols = a.value_cols;
		p.aggs = a.aggs;
		p.seg_out = a.seg_out;
		p.key_out = a.key_out;
		p.agg_out = a.agg_out;
		p.capacity = a.capacity;
		auto h = std::make_unique<vgpu_batched_plan>();
		h->plan = std::make_unique<BatchedGraphPlan>(p);
		*plan = h.release();
	});
}

int vgpu_batched_plan_run(vgpu_batched_plan *plan, const uint64_t *row_starts, const uint64_t *row_lens,
                          int32_t nranges, uint64_t *groups, uint64_t *segment_group_offsets, void *stream) {
	return Guard([&] {
		if (plan == nullptr || plan->plan == nullptr) {
			throw std::runtime_error("batched plan: NULL plan");
		}
		if (nranges < 0 || (nranges > 0 && (row_starts == nullptr || row_lens == nullptr))) {
			throw std::runtime_error("batched plan: bad row ranges");
		}
		*groups = plan->plan->Run(row_starts, row_lens, (uint32_t)nranges, segment_group_offsets,
		                          static_cast<hipStream_t>(stream));
	});
}

int vgpu_batched_plan_info(const vgpu_batched_plan *plan, uint64_t *kernel_nodes, uint64_t *copy_nodes,
                           uint64_t *other_nodes, uint64_t *footprint_bytes, uint64_t *group_bound) {
	return Guard([&] {
		if (plan == nullptr || plan->plan == nullptr) {
			throw std::runtime_error("batched plan: NULL plan");
		}
		const auto &p = *plan->plan;
		if (kernel_nodes != nullptr) {
			*kernel_nodes = p.kernel_nodes();
		}
		if (copy_nodes != nullptr) {
			*copy_nodes = p.copy_nodes();
		}
		if (other_nodes != nullptr) {
			*other_nodes = p.other_nodes();
		}
		if (footprint_bytes != nullptr) {
			*footprint_bytes = p.footprint();
		}
		if (group_bound != nullptr) {
			*group_bound = p.group_bound();
		}
	});
}

void vgpu_batched_plan_destroy(vgpu_batched_plan *plan) {
	delete plan;
}

} // extern "C"
