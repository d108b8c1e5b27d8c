// gpu_stream_pipeline.cpp -- GpuStreamPipeline on three non-blocking HIP streams (contract of
// cuda_engine/include/gpu_stream_pipeline.hpp and cuda_engine/src/gpu_stream_pipeline.cu:347-837:
// kDepth batches in flight, kStageCount stage sets, FIFO results, input slot released when its upload has
// completed, the slot accounting that keeps the producer from deadlocking, RUNNING_SUM carried across
// batches, schema and size checks with the same messages).
//
// Differences in execution: each projection level is one generated kernel for all of its expressions
// (reading each input column once, writing each output once) instead of one kernel per expression; NULL
// validity of the top level is packed into the Arrow bitmap on the device, so the host does no packing.
#include "../../../include/code_generator.hpp"
#include "../../../include/gpu_logger.hpp"
#include "../../../include/gpu_stream_pipeline.hpp"
#include "../hip/kernels.h"
#include "fused.h"

#include <hip/hip_runtime.h>

#include <array>
#include <deque>
#include <map>
#include <set>
#include <stdexcept>
#include <string>

namespace vector_gpu {

static_assert(GpuStreamPipeline::kDepth + 1 <= GpuMemoryPool::kSlotCount,
              "the producer needs a free pool slot while kDepth batches hold theirs");

namespace {

void Hip(hipError_t status, const char *what) {
	if (status != hipSuccess) {
		(void)hipGetLastError();
		throw std::runtime_error(std::string("gpu_stream_pipeline: ") + what + " failed: " + hipGetErrorString(status));
	}
}

std::string OutputName(size_t i) {
	return "col" + std::to_string(i);
}

bool FusableType(GpuValueType t) {
	switch (t) {
	case GpuValueType::INT16:
	case GpuValueType::INT32:
	case GpuValueType::INT64:
	case GpuValueType::FLOAT32:
	case GpuValueType::FLOAT64:
	case GpuValueType::BOOLEAN:
		return true;
	default:
		return false;
	}
}

bool RunningSumType(GpuValueType t) {
	return FusableType(t) && t != GpuValueType::BOOLEAN;
}

size_t Width(GpuValueType t) {
	return static_cast<size_t>(vgpu::TypeWidth(static_cast<int>(t)));
}

//! The PROJECTION chain from `plan` down to its SCAN, bottom-up; nullptr when the plan is not that shape.
const GpuPlanNode *ProjectionChain(const GpuPlanNode &plan, std::vector<const GpuPlanNode *> &levels) {
	std::vector<const GpuPlanNode *> top_down;
	const GpuPlanNode *node = &plan;
	while (node->op_type == GpuOpType::PROJECTION) {
		if (node->children.size() != 1 || !node->children[0] || node->expressions.empty()) {
			return nullptr;
		}
		top_down.push_back(node);
		node = node->children[0].get();
	}
	if (top_down.empty() || node->op_type != GpuOpType::SCAN) {
		return nullptr;
	}
	levels.assign(top_down.rbegin(), top_down.rend());
	return node;
}

std::vector<std::string> SourceNames(const std::vector<std::string> &inputs,
                                     const std::vector<const GpuPlanNode *> &levels, size_t level) {
	if (level == 0) {
		return inputs;
	}
	std::vector<std::string> names;
	for (size_t i = 0; i < levels[level - 1]->expressions.size(); i++) {
		names.push_back(OutputName(i));
	}
	return names;
}

bool Resolve(const GpuExpr &expr, const std::vector<std::string> &sources, std::vector<int> &out) {
	out.clear();
	for (auto &name : expr.input_columns) {
		int found = -1;
		for (size_t i = 0; i < sources.size(); i++) {
			if (sources[i] == name) {
				found = static_cast<int>(i);
				break;
			}
		}
		if (found < 0) {
			return false;
		}
		out.push_back(found);
	}
	return true;
}

} // namespace

//! Expressions of one level evaluated by one generated kernel (a level splits only past 16 columns or
//! outputs). `plain` runs batches without NULLs, `nullable` batches where some input carries them.
struct PipelineKernelGroup {
	std::vector<int> exprs;
	std::vector<int> sources; // group slot -> level source index
	std::shared_ptr<vgpu::FusedKernel> plain;
	std::shared_ptr<vgpu::FusedKernel> nullable;
};

struct PipelineLevelSpec {
	std::vector<GpuExpr> exprs;
	std::vector<GpuValueType> out_types;
	std::vector<std::vector<int>> inputs;
	std::vector<PipelineKernelGroup> groups;
};

struct PipelineStage {
	std::vector<void *> d_in;
	std::vector<uint8_t *> d_valid;
	std::vector<uint8_t *> d_packed;
	std::vector<std::vector<void *>> d_out;         // per level, per expression
	std::vector<std::vector<uint8_t *>> d_out_valid; // per level, per expression
	std::vector<uint8_t *> d_top_packed;             // top level, per expression
	uint8_t *h_out = nullptr;                        // page-locked: top data regions, then packed validity
	hipEvent_t h2d_done = nullptr;
	hipEvent_t exec_done = nullptr;
	hipEvent_t d2h_done = nullptr;
	uint64_t rows = 0;
	bool any_null = false;
	GpuMemoryPoolSlot input_slot;
	bool input_released = true;
};

struct GpuStreamPipelineImpl {
	std::vector<std::string> input_names;
	std::vector<GpuValueType> input_types;
	std::vector<PipelineLevelSpec> levels;
	std::vector<size_t> out_data_offset;
	std::vector<size_t> out_valid_offset;
	size_t h_out_bytes = 0;
	uint64_t max_rows = 0;
	hipStream_t stream_h2d = nullptr;
	hipStream_t stream_exec = nullptr;
	hipStream_t stream_d2h = nullptr;
	std::array<PipelineStage, GpuStreamPipeline::kStageCount> stages;
	std::deque<int> in_flight;
	std::vector<int> free_stages;
	size_t input_slots_held = 0;
	bool failed = false;
	std::vector<void *> device_allocs;
	std::vector<std::vector<void *>> running_state; // per level, per expression (nullptr if none)
	void *scan_scratch = nullptr;
	size_t scan_scratch_bytes = 0;

	const PipelineLevelSpec &top() const {
		return levels.back();
	}

	void *AllocDevice(size_t bytes) {
		void *p = nullptr;
		Hip(hipMalloc(&p, bytes == 0 ? 1 : bytes), "hipMalloc of a stage buffer");
		device_allocs.push_back(p);
		return p;
	}

	~GpuStreamPipelineImpl() {
		for (auto s : {stream_h2d, stream_exec, stream_d2h}) {
			if (s != nullptr) {
				(void)hipStreamSynchronize(s);
			}
		}
		for (auto &stage : stages) {
			for (auto e : {stage.h2d_done, stage.exec_done, stage.d2h_done}) {
				if (e != nullptr) {
					(void)hipEventDestroy(e);
				}
			}
			if (stage.h_out != nullptr) {
				(void)hipHostFree(stage.h_out);
			}
		}
		for (void *p : device_allocs) {
			(void)hipFree(p);
		}
		for (auto s : {stream_h2d, stream_exec, stream_d2h}) {
			if (s != nullptr) {
				(void)hipStreamDestroy(s);
			}
		}
	}
};

// ---------------------------------------------------------------- GpuStreamResult

GpuStreamResult::GpuStreamResult(GpuStreamResult &&other) noexcept
    : pipeline_(other.pipeline_), stage_(other.stage_), columns_(std::move(other.columns_)),
      row_count_(other.row_count_) {
	other.pipeline_ = nullptr;
	other.stage_ = -1;
	other.row_count_ = 0;
}

GpuStreamResult &GpuStreamResult::operator=(GpuStreamResult &&other) noexcept {
	if (this != &other) {
		Reset();
		pipeline_ = other.pipeline_;
		stage_ = other.stage_;
		columns_ = std::move(other.columns_);
		row_count_ = other.row_count_;
		other.pipeline_ = nullptr;
		other.stage_ = -1;
		other.row_count_ = 0;
	}
	return *this;
}

GpuStreamResult::~GpuStreamResult() {
	Reset();
}

void GpuStreamResult::Reset() {
	if (pipeline_ != nullptr) {
		pipeline_->ReleaseStage(stage_);
	}
	pipeline_ = nullptr;
	stage_ = -1;
	columns_.clear();
	row_count_ = 0;
}

// ---------------------------------------------------------------- GpuStreamPipeline

bool GpuStreamPipeline::Supports(const GpuPlanNode &plan) {
	std::vector<const GpuPlanNode *> levels;
	const GpuPlanNode *scan = ProjectionChain(plan, levels);
	if (scan == nullptr || scan->column_names.empty()) {
		return false;
	}
	std::vector<int> idx;
	for (size_t level = 0; level < levels.size(); level++) {
		auto sources = SourceNames(scan->column_names, levels, level);
		for (auto &e : levels[level]->expressions) {
			if (!FusableType(e.output_type) || e.generated_cuda_source.empty() || !Resolve(e, sources, idx)) {
				return false;
			}
			if (e.window_primitive == GpuWindowPrimitive::RUNNING_SUM && !RunningSumType(e.output_type)) {
				return false;
			}
		}
	}
	return true;
}

GpuStreamPipeline::GpuStreamPipeline(const GpuPlanNode &plan, const std::vector<std::string> &input_names,
                                     const std::vector<GpuValueType> &input_types, uint64_t max_rows_per_batch)
    : impl_(new GpuStreamPipelineImpl()) {
	std::vector<const GpuPlanNode *> plan_levels;
	if (ProjectionChain(plan, plan_levels) == nullptr) {
		throw std::runtime_error("gpu_stream_pipeline: plan is not a chain of PROJECTIONs over a SCAN");
	}
	if (input_names.size() != input_types.size() || input_names.empty()) {
		throw std::runtime_error("gpu_stream_pipeline: input_names and input_types must be the same non-empty "
		                         "length");
	}
	if (max_rows_per_batch == 0) {
		throw std::runtime_error("gpu_stream_pipeline: max_rows_per_batch must be > 0");
	}
	for (auto t : input_types) {
		if (!FusableType(t)) {
			throw std::runtime_error("gpu_stream_pipeline: input type has no representation a generated kernel "
			                         "reads");
		}
	}
	auto &impl = *impl_;
	impl.max_rows = max_rows_per_batch;
	impl.input_names = input_names;
	impl.input_types = input_types;
	const bool fast = IsGpuFastMathEnabled();

	// Levels, their kernel groups, and both kernel variants, compiled now.
	std::vector<GpuValueType> source_types = input_types;
	for (size_t level = 0; level < plan_levels.size(); level++) {
		auto sources = SourceNames(input_names, plan_levels, level);
		PipelineLevelSpec spec;
		for (auto &e : plan_levels[level]->expressions) {
			std::vector<int> idx;
			if (!Resolve(e, sources, idx)) {
				throw std::runtime_error("gpu_stream_pipeline: a projection expression references a column its "
				                         "input does not produce");
			}
			if (!FusableType(e.output_type)) {
				throw std::runtime_error("gpu_stream_pipeline: expression output type is not supported");
			}
			if (e.window_primitive == GpuWindowPrimitive::RUNNING_SUM && !RunningSumType(e.output_type)) {
				throw std::runtime_error("gpu_stream_pipeline: unsupported type for RUNNING_SUM");
			}
			spec.exprs.push_back(e);
			spec.out_types.push_back(e.output_type);
			spec.inputs.push_back(idx);
		}
		for (size_t e = 0; e < spec.exprs.size();) {
			PipelineKernelGroup g;
			std::map<int, int> slot_of;
			for (; e < spec.exprs.size() && g.exprs.size() < (size_t)vgpu::kFusedMaxOuts; e++) {
				std::set<int> fresh;
				for (int s : spec.inputs[e]) {
					if (!slot_of.count(s)) {
						fresh.insert(s);
					}
				}
				if (g.sources.size() + fresh.size() > (size_t)vgpu::kFusedMaxCols) {
					if (g.exprs.empty()) {
						throw std::runtime_error("gpu_stream_pipeline: an expression reads more than 16 columns");
					}
					break;
				}
				for (int s : spec.inputs[e]) {
					if (!slot_of.count(s)) {
						slot_of.emplace(s, static_cast<int>(g.sources.size()));
						g.sources.push_back(s);
					}
				}
				g.exprs.push_back(static_cast<int>(e));
			}
			for (int variant = 0; variant < 2; variant++) {
				vgpu::FusedSpec fs;
				fs.fast_math = fast;
				for (int s : g.sources) {
					vgpu::FusedColumn fc;
					fc.vtype = static_cast<int>(source_types[static_cast<size_t>(s)]);
					fc.has_valid = variant == 1;
					fs.cols.push_back(fc);
				}
				for (int x : g.exprs) {
					vgpu::FusedExpr fe;
					fe.body = spec.exprs[static_cast<size_t>(x)].generated_cuda_source;
					fe.out_type = static_cast<int>(spec.out_types[static_cast<size_t>(x)]);
					for (int s : spec.inputs[static_cast<size_t>(x)]) {
						fe.slots.push_back(slot_of.at(s));
					}
					fs.projections.push_back(fe);
				}
				(variant == 0 ? g.plain : g.nullable) = vgpu::GetFusedKernel(fs);
			}
			spec.groups.push_back(std::move(g));
		}
		source_types = spec.out_types;
		impl.levels.push_back(std::move(spec));
	}

	size_t offset = 0;
	for (auto t : impl.top().out_types) {
		impl.out_data_offset.push_back(offset);
		offset += (static_cast<size_t>(max_rows_per_batch) * Width(t) + 255) / 256 * 256;
	}
	for (size_t e = 0; e < impl.top().out_types.size(); e++) {
		impl.out_valid_offset.push_back(offset);
		offset += ((static_cast<size_t>(max_rows_per_batch) + 7) / 8 + 255) / 256 * 256;
	}
	impl.h_out_bytes = offset;

	Hip(hipStreamCreateWithFlags(&impl.stream_h2d, hipStreamNonBlocking), "hipStreamCreate(h2d)");
	Hip(hipStreamCreateWithFlags(&impl.stream_exec, hipStreamNonBlocking), "hipStreamCreate(exec)");
	Hip(hipStreamCreateWithFlags(&impl.stream_d2h, hipStreamNonBlocking), "hipStreamCreate(d2h)");
	const size_t rows = static_cast<size_t>(max_rows_per_batch);
	for (size_t s = 0; s < kStageCount; s++) {
		auto &st = impl.stages[s];
		Hip(hipEventCreateWithFlags(&st.h2d_done, hipEventDisableTiming), "hipEventCreate");
		Hip(hipEventCreateWithFlags(&st.exec_done, hipEventDisableTiming), "hipEventCreate");
		Hip(hipEventCreateWithFlags(&st.d2h_done, hipEventDisableTiming), "hipEventCreate");
		for (auto t : input_types) {
			st.d_in.push_back(impl.AllocDevice(rows * Width(t)));
			st.d_valid.push_back(static_cast<uint8_t *>(impl.AllocDevice(rows)));
			st.d_packed.push_back(static_cast<uint8_t *>(impl.AllocDevice((rows + 7) / 8)));
		}
		for (auto &level : impl.levels) {
			std::vector<void *> outs;
			std::vector<uint8_t *> valids;
			for (auto t : level.out_types) {
				outs.push_back(impl.AllocDevice(rows * Width(t)));
				valids.push_back(static_cast<uint8_t *>(impl.AllocDevice(rows)));
			}
			st.d_out.push_back(outs);
			st.d_out_valid.push_back(valids);
		}
		for (size_t e = 0; e < impl.top().out_types.size(); e++) {
			st.d_top_packed.push_back(static_cast<uint8_t *>(impl.AllocDevice((rows + 7) / 8)));
		}
		Hip(hipHostMalloc(reinterpret_cast<void **>(&st.h_out), impl.h_out_bytes, hipHostMallocPortable),
		    "hipHostMalloc of a stage's result buffer");
		impl.free_stages.push_back(static_cast<int>(s));
	}
	bool any_running = false;
	for (auto &level : impl.levels) {
		std::vector<void *> states;
		for (size_t e = 0; e < level.exprs.size(); e++) {
			if (level.exprs[e].window_primitive == GpuWindowPrimitive::RUNNING_SUM) {
				void *p = impl.AllocDevice(16);
				vgpu::FillU64(static_cast<uint64_t *>(p), 0ull, 2, impl.stream_exec);
				states.push_back(p);
				any_running = true;
			} else {
				states.push_back(nullptr);
			}
		}
		impl.running_state.push_back(states);
	}
	if (any_running) {
		impl.scan_scratch_bytes = vgpu::RunningSumScratchBytes(max_rows_per_batch);
		impl.scan_scratch = impl.AllocDevice(impl.scan_scratch_bytes);
	}
	Hip(hipStreamSynchronize(impl.stream_exec), "initial state fill");
	VGPU_LOG(LogLevel::INFO, "stream_pipeline_init",
	         "\"depth\":" + std::to_string(kDepth) + ",\"stages\":" + std::to_string(kStageCount) + ",\"max_rows\":" +
	             std::to_string(max_rows_per_batch) + ",\"inputs\":" + std::to_string(input_names.size()) +
	             ",\"levels\":" + std::to_string(impl.levels.size()) + ",\"outputs\":" +
	             std::to_string(impl.top().exprs.size()));
}

GpuStreamPipeline::~GpuStreamPipeline() = default;

const std::vector<GpuValueType> &GpuStreamPipeline::output_types() const {
	return impl_->top().out_types;
}

size_t GpuStreamPipeline::in_flight() const {
	return impl_->in_flight.size();
}

void GpuStreamPipeline::ReleaseStage(int stage) {
	impl_->free_stages.push_back(stage);
}

void GpuStreamPipeline::Poll() {
	auto &impl = *impl_;
	for (int s : impl.in_flight) {
		auto &st = impl.stages[static_cast<size_t>(s)];
		if (st.input_released) {
			continue;
		}
		hipError_t status = hipEventQuery(st.h2d_done);
		if (status == hipErrorNotReady) {
			continue;
		}
		if (status != hipSuccess) {
			(void)hipGetLastError();
			impl.failed = true;
			throw std::runtime_error(std::string("gpu_stream_pipeline: hipEventQuery on the H2D event failed: ") +
			                         hipGetErrorString(status));
		}
		st.input_slot = GpuMemoryPoolSlot();
		st.input_released = true;
		impl.input_slots_held--;
	}
}

bool GpuStreamPipeline::CanSubmit() {
	Poll();
	auto &impl = *impl_;
	if (impl.failed || impl.in_flight.size() >= kDepth || impl.free_stages.empty()) {
		return false;
	}
	// The producer must keep one pool slot to fill its next batch into (it blocks on the ring otherwise).
	return impl.input_slots_held + 2 <= GpuMemoryPool::kSlotCount;
}

void GpuStreamPipeline::Submit(GpuStreamBatch batch) {
	auto &impl = *impl_;
	if (!CanSubmit()) {
		throw std::runtime_error("gpu_stream_pipeline: Submit called when the pipeline cannot accept a batch "
		                         "(check CanSubmit() first)");
	}
	if (batch.row_count == 0) {
		throw std::runtime_error("gpu_stream_pipeline: refusing an empty batch");
	}
	if (batch.row_count > impl.max_rows) {
		throw std::runtime_error("gpu_stream_pipeline: batch of " + std::to_string(batch.row_count) +
		                         " rows exceeds the " + std::to_string(impl.max_rows) +
		                         "-row size every stage buffer was allocated for");
	}
	if (batch.columns.size() != impl.input_names.size()) {
		throw std::runtime_error("gpu_stream_pipeline: batch has " + std::to_string(batch.columns.size()) +
		                         " columns but the declared input schema has " +
		                         std::to_string(impl.input_names.size()));
	}
	const int si = impl.free_stages.back();
	auto &st = impl.stages[static_cast<size_t>(si)];
	bool any_null = false;
	for (size_t i = 0; i < batch.columns.size(); i++) {
		auto &c = batch.columns[i];
		if (c.name != impl.input_names[i] || c.type != impl.input_types[i]) {
			throw std::runtime_error("gpu_stream_pipeline: batch column " + std::to_string(i) + " ('" + c.name +
			                         "') does not match the declared schema entry '" + impl.input_names[i] + "'");
		}
		if (c.row_count != batch.row_count) {
			throw std::runtime_error("gpu_stream_pipeline: column '" + c.name + "' carries " +
			                         std::to_string(c.row_count) + " rows but the batch declares " +
			                         std::to_string(batch.row_count));
		}
		any_null = any_null || c.validity != nullptr;
	}
	const uint64_t rows = batch.row_count;
	try {
		for (size_t i = 0; i < batch.columns.size(); i++) {
			auto &c = batch.columns[i];
			Hip(hipMemcpyAsync(st.d_in[i], c.data, rows * Width(c.type), hipMemcpyHostToDevice, impl.stream_h2d),
			    "H2D upload of a batch column");
			if (c.validity != nullptr) {
				Hip(hipMemcpyAsync(st.d_packed[i], c.validity, (rows + 7) / 8, hipMemcpyHostToDevice, impl.stream_h2d),
				    "H2D upload of a validity bitmap");
			}
		}
		Hip(hipEventRecord(st.h2d_done, impl.stream_h2d), "hipEventRecord(h2d)");
		Hip(hipStreamWaitEvent(impl.stream_exec, st.h2d_done, 0), "hipStreamWaitEvent(h2d)");
		if (any_null) {
			for (size_t i = 0; i < batch.columns.size(); i++) {
				if (batch.columns[i].validity != nullptr) {
					vgpu::UnpackValidity(st.d_packed[i], st.d_valid[i], rows, impl.stream_exec);
				} else {
					vgpu::FillU8(st.d_valid[i], 1, rows, impl.stream_exec);
				}
			}
		}
		for (size_t level = 0; level < impl.levels.size(); level++) {
			auto &spec = impl.levels[level];
			const std::vector<void *> &src = level == 0 ? st.d_in : st.d_out[level - 1];
			const std::vector<uint8_t *> &src_valid = level == 0 ? st.d_valid : st.d_out_valid[level - 1];
			for (auto &g : spec.groups) {
				std::vector<const void *> cols;
				std::vector<const uint8_t *> valid;
				for (int s : g.sources) {
					cols.push_back(src[static_cast<size_t>(s)]);
					valid.push_back(any_null ? src_valid[static_cast<size_t>(s)] : nullptr);
				}
				std::vector<void *> outs;
				std::vector<uint8_t *> out_valid;
				for (int x : g.exprs) {
					outs.push_back(st.d_out[level][static_cast<size_t>(x)]);
					out_valid.push_back(any_null ? st.d_out_valid[level][static_cast<size_t>(x)] : nullptr);
					if (any_null && spec.inputs[static_cast<size_t>(x)].empty()) {
						// A constant expression is never NULL; the kernel writes no validity for it.
						vgpu::FillU8(st.d_out_valid[level][static_cast<size_t>(x)], 1, rows, impl.stream_exec);
					}
				}
				vgpu::FusedLaunch la;
				la.cols = cols.data();
				la.valid = valid.data();
				la.outs = outs.data();
				la.out_valid = out_valid.data();
				la.n = rows;
				(any_null ? g.nullable : g.plain)->Launch(la, nullptr, 0, impl.stream_exec);
			}
			for (size_t e = 0; e < spec.exprs.size(); e++) {
				if (impl.running_state[level][e] != nullptr) {
					vgpu::RunningSum(static_cast<int>(spec.out_types[e]), st.d_out[level][e], rows,
					                 impl.running_state[level][e], impl.scan_scratch, impl.scan_scratch_bytes,
					                 impl.stream_exec);
				}
			}
		}
		const size_t top = impl.levels.size() - 1;
		if (any_null) {
			for (size_t e = 0; e < impl.top().exprs.size(); e++) {
				vgpu::PackValidity(st.d_out_valid[top][e], st.d_top_packed[e], rows, impl.stream_exec);
			}
		}
		Hip(hipEventRecord(st.exec_done, impl.stream_exec), "hipEventRecord(exec)");
		Hip(hipStreamWaitEvent(impl.stream_d2h, st.exec_done, 0), "hipStreamWaitEvent(exec)");
		for (size_t e = 0; e < impl.top().exprs.size(); e++) {
			Hip(hipMemcpyAsync(st.h_out + impl.out_data_offset[e], st.d_out[top][e], rows * Width(impl.top().out_types[e]),
			                   hipMemcpyDeviceToHost, impl.stream_d2h),
			    "D2H copy of a result column");
			if (any_null) {
				Hip(hipMemcpyAsync(st.h_out + impl.out_valid_offset[e], st.d_top_packed[e], (rows + 7) / 8,
				                   hipMemcpyDeviceToHost, impl.stream_d2h),
				    "D2H copy of a result column's validity");
			}
		}
		Hip(hipEventRecord(st.d2h_done, impl.stream_d2h), "hipEventRecord(d2h)");
	} catch (...) {
		impl.failed = true;
		throw;
	}
	impl.free_stages.pop_back();
	st.rows = rows;
	st.any_null = any_null;
	st.input_slot = std::move(batch.slot);
	st.input_released = false;
	impl.input_slots_held++;
	impl.in_flight.push_back(si);
	VGPU_LOG(LogLevel::DEBUG, "stream_submit",
	         "\"stage\":" + std::to_string(si) + ",\"rows\":" + std::to_string(rows) + ",\"in_flight\":" +
	             std::to_string(impl.in_flight.size()) + ",\"nullable\":" + (any_null ? "true" : "false"));
}

namespace {
void TakeStage(GpuStreamPipelineImpl &impl, int si, std::vector<GpuColumn> &columns, uint64_t &row_count) {
	auto &st = impl.stages[static_cast<size_t>(si)];
	columns.clear();
	for (size_t e = 0; e < impl.top().exprs.size(); e++) {
		GpuColumn c;
		c.name = OutputName(e);
		c.type = impl.top().out_types[e];
		c.data = st.h_out + impl.out_data_offset[e];
		c.row_count = st.rows;
		c.validity = st.any_null ? st.h_out + impl.out_valid_offset[e] : nullptr;
		columns.push_back(std::move(c));
	}
	row_count = st.rows;
	if (!st.input_released) {
		st.input_slot = GpuMemoryPoolSlot();
		st.input_released = true;
		impl.input_slots_held--;
	}
	impl.in_flight.pop_front();
}
} // namespace

bool GpuStreamPipeline::TryTake(GpuStreamResult &out) {
	auto &impl = *impl_;
	Poll();
	if (impl.in_flight.empty()) {
		return false;
	}
	const int si = impl.in_flight.front();
	hipError_t status = hipEventQuery(impl.stages[static_cast<size_t>(si)].d2h_done);
	if (status == hipErrorNotReady) {
		return false;
	}
	if (status != hipSuccess) {
		(void)hipGetLastError();
		impl.failed = true;
		throw std::runtime_error(std::string("gpu_stream_pipeline: hipEventQuery on the D2H event failed: ") +
		                         hipGetErrorString(status));
	}
	out.Reset();
	TakeStage(impl, si, out.columns_, out.row_count_);
	out.pipeline_ = this;
	out.stage_ = si;
	return true;
}

bool GpuStreamPipeline::TakeBlocking(GpuStreamResult &out) {
	auto &impl = *impl_;
	if (impl.in_flight.empty()) {
		return false;
	}
	const int si = impl.in_flight.front();
	hipError_t status = hipEventSynchronize(impl.stages[static_cast<size_t>(si)].d2h_done);
	if (status != hipSuccess) {
		(void)hipGetLastError();
		impl.failed = true;
		throw std::runtime_error(std::string("gpu_stream_pipeline: hipEventSynchronize on the D2H event failed: ") +
		                         hipGetErrorString(status));
	}
	out.Reset();
	TakeStage(impl, si, out.columns_, out.row_count_);
	out.pipeline_ = this;
	out.stage_ = si;
	return true;
}

} // namespace vector_gpu
