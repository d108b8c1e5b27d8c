// batched_graph.h -- the batched filter + GROUP BY form prepared once and run as one HIP graph per batch.
#pragma once

#include "operators.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

namespace vgpu {

//! What a plan fixes when it is prepared. Each batch is up to max_queries row ranges of the table
//! (range g = table rows [start[g], start[g] + len[g]); ranges may overlap or repeat) holding at most
//! max_rows rows together; every range runs the same predicates (slots index `cols`) and groups by
//! `key_col` with `aggs` (inputs index `value_cols`). Outputs are the caller's device columns of `capacity`
//! rows (segment INT32, key, one per aggregate), which must stay valid while the plan lives; groups are
//! ordered by (range, key).
struct BatchedPlanArgs {
	std::vector<DevCol> cols;
	uint64_t table_rows = 0;
	uint32_t max_queries = 0;
	uint64_t max_rows = 0;
	std::vector<FusedExpr> preds;
	int key_col = -1;
	std::vector<int> value_cols;
	std::vector<AggReq> aggs;
	void *seg_out = nullptr;
	void *key_out = nullptr;
	std::vector<void *> agg_out;
	uint64_t capacity = 0;
	bool fast_math = false;
};

class BatchedGraphPlan {
public:
	//! Compiles the filter, reads the key column's range, sizes and allocates every buffer for the bounds,
	//! runs the pipeline once on an empty batch and captures it into a graph. Throws when the bounds do not
	//! fit (capacity below the group bound, rows past 2^30).
	explicit BatchedGraphPlan(const BatchedPlanArgs &args);
	~BatchedGraphPlan();
	BatchedGraphPlan(const BatchedGraphPlan &) = delete;
	BatchedGraphPlan &operator=(const BatchedGraphPlan &) = delete;

	//! Runs one batch of nq ranges: checks them against the table and the bounds (before any GPU work),
	//! launches the graph on `stream` and waits for it. Returns the group count; seg_group_offsets (nq + 1
	//! entries, optional) receives each range's first group. Calls on one plan must not overlap.
	uint64_t Run(const uint64_t *row_starts, const uint64_t *row_lens, uint32_t nq, uint64_t *seg_group_offsets,
	             hipStream_t stream);

	size_t kernel_nodes() const {
		return kernel_nodes_;
	}
	size_t copy_nodes() const {
		return copy_nodes_;
	}
	size_t other_nodes() const {
		return other_nodes_;
	}
	//! Device bytes the plan holds (its buffers and the pipeline's scratch, all sized for the bounds).
	size_t footprint() const {
		return footprint_;
	}
	uint64_t group_bound() const {
		return group_bound_;
	}

private:
	class ReplayScratch;
	void Prepare();
	void Enqueue(hipStream_t stream);
	void Release();

	BatchedPlanArgs args_;
	std::shared_ptr<FusedKernel> kernel_;
	std::vector<const void *> col_ptrs_;
	KeyPack pack_;
	uint64_t key_min_ = 0, key_max_ = 0;
	uint64_t group_bound_ = 0;
	hipStream_t stream_ = nullptr;
	std::unique_ptr<ReplayScratch> scratch_;
	uint64_t *queue_host_ = nullptr;
	uint64_t *result_host_ = nullptr;
	uint64_t queue_words_ = 0;
	uint64_t result_words_ = 0;
	hipGraph_t graph_ = nullptr;
	hipGraphExec_t exec_ = nullptr;
	size_t kernel_nodes_ = 0, copy_nodes_ = 0, other_nodes_ = 0;
	size_t footprint_ = 0;
};

} // namespace vgpu
