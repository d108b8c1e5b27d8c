#pragma once
// gpu_stream_pipeline.hpp -- three-stream pipeline for a chain of projections over streamed batches
// (drop-in for cuda_engine/include/gpu_stream_pipeline.hpp).
//
// Batch N+2 uploads, batch N+1 computes and batch N downloads at the same time, each on its own
// non-blocking HIP stream (default priority), ordered only by events. The supported shape is the
// reference's: one or more PROJECTIONs ending in a SCAN (output row count known before the kernels run),
// so every device buffer is sized once, in the constructor, and no allocation happens per batch. Each
// projection level runs as ONE generated gfx950 kernel for all of its expressions (split only when a level
// reads or writes more than 16 columns); a RUNNING_SUM expression is followed by a carried running sum.
// Results come back in submission order into page-locked stage buffers owned by the pipeline.
//
// Externally serialized: calls must not overlap (no locks inside; the calling thread never blocks except
// in TakeBlocking).

#include "gpu_engine.hpp"
#include "gpu_memory_pool.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace vector_gpu {

struct GpuStreamPipelineImpl;

//! One batch of page-locked host rows. `slot` owns the memory every column points into; Submit() takes it
//! and returns it to GpuMemoryPool as soon as the batch's upload has completed. `columns` follow the
//! pipeline's declared input schema (names and types are checked).
struct GpuStreamBatch {
	GpuMemoryPoolSlot slot;
	std::vector<GpuColumn> columns;
	uint64_t row_count = 0;
};

//! One completed batch's results in page-locked memory owned by the pipeline stage that produced them.
//! The stage returns to the pipeline when this object is destroyed, overwritten or Reset(); the column
//! pointers dangle from then on.
class GpuStreamResult {
public:
	GpuStreamResult() = default;
	GpuStreamResult(const GpuStreamResult &) = delete;
	GpuStreamResult &operator=(const GpuStreamResult &) = delete;
	GpuStreamResult(GpuStreamResult &&other) noexcept;
	GpuStreamResult &operator=(GpuStreamResult &&other) noexcept;
	~GpuStreamResult();

	//! One column per expression of the top projection level, named col0, col1, ...
	const std::vector<GpuColumn> &columns() const {
		return columns_;
	}
	uint64_t row_count() const {
		return row_count_;
	}
	bool valid() const {
		return pipeline_ != nullptr;
	}
	void Reset();

private:
	friend class GpuStreamPipeline;
	class GpuStreamPipeline *pipeline_ = nullptr;
	int stage_ = -1;
	std::vector<GpuColumn> columns_;
	uint64_t row_count_ = 0;
};

class GpuStreamPipeline {
public:
	//! Batches in flight at once (one uploading, one computing, one downloading).
	static constexpr size_t kDepth = 3;
	//! Stage buffer sets: one more than kDepth for the result the consumer is holding.
	static constexpr size_t kStageCount = kDepth + 1;

	//! True for a chain of one or more PROJECTIONs (each with at least one expression) ending in a SCAN,
	//! every expression's inputs resolvable against the level below (the SCAN's columns, then col<i>),
	//! and every output type one a generated kernel can write (fixed width, not HUGEINT).
	static bool Supports(const GpuPlanNode &plan);

	//! Allocates every device and page-locked buffer, the three streams and the per-stage events, and
	//! compiles every level's kernels, before any row is accepted. Throws std::runtime_error when the plan
	//! is unsupported, the schema is inconsistent, or a resource cannot be acquired.
	GpuStreamPipeline(const GpuPlanNode &plan, const std::vector<std::string> &input_names,
	                  const std::vector<GpuValueType> &input_types, uint64_t max_rows_per_batch);
	~GpuStreamPipeline();

	GpuStreamPipeline(const GpuStreamPipeline &) = delete;
	GpuStreamPipeline &operator=(const GpuStreamPipeline &) = delete;

	const std::vector<GpuValueType> &output_types() const;

	//! Releases the input slot of every in-flight batch whose upload has completed (event queries only).
	void Poll();

	//! True when Submit() would neither exceed kDepth batches in flight, nor find no free stage, nor leave
	//! the producer without a free GpuMemoryPool slot for its next batch. Calls Poll().
	bool CanSubmit();

	//! Queues upload, kernels and download of one batch; returns once queued. Throws std::runtime_error
	//! when !CanSubmit(), the batch is empty, larger than max_rows_per_batch, or off-schema, or a HIP call
	//! fails.
	void Submit(GpuStreamBatch batch);

	//! The oldest in-flight batch's results if its download has finished (FIFO); false otherwise.
	bool TryTake(GpuStreamResult &out);
	//! Same, waiting for the oldest batch; false only when nothing is in flight.
	bool TakeBlocking(GpuStreamResult &out);

	size_t in_flight() const;

private:
	friend class GpuStreamResult;
	void ReleaseStage(int stage);
	std::unique_ptr<GpuStreamPipelineImpl> impl_;
};

} // namespace vector_gpu
