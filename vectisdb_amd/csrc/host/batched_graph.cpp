// batched_graph.cpp -- the batched filter + GROUP BY form prepared once and replayed as one HIP graph.
//
// BatchedFilterGroupByDevice plans each batch on the host: it reads the kept-row count and the kept keys'
// ranges back after the filter and sizes the GROUP BY from them, so a batch costs about ten launches and
// three host round trips. A prepared plan fixes everything the host decides up front, from bounds that
// hold for every batch it will run:
//   * the composite key (segment, key) is packed from the segment count bound and the key column's
//     table-wide range (read once, at preparation), not from the kept rows;
//   * every buffer is sized for the row bound (at most max_rows rows over at most max_queries row ranges),
//     and every kernel after the filter reads the kept-row count from device memory;
//   * the filter is the persistent form of the generated kernel: a fixed grid of resident blocks sweeps
//     the tiles of a work queue built on the device from the batch's row-range descriptors.
// The whole pipeline -- descriptor upload, queue layout, look-back reset, filter, radix passes, segmented
// reduction, per-query group offsets, result download -- is captured once into a graph. Running a batch
// writes its descriptors into pinned memory, launches the graph and waits for it: one launch per batch,
// no host read or synchronizing call inside it. Results equal BatchedFilterGroupByDevice's over the same
// row ranges (same compaction order, same sort, same reduction tiles).
#include "batched_graph.h"

#include "device_pool.h"
#include "../hip/common.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>

namespace vgpu {

namespace {

constexpr int kLayoutBlock = 1024;

//! Queue layout of one batch (one block). queue = {nq, first table row of each range (qmax words), rows of
//! each range (qmax words)}. Range g becomes launch rows [seg[g], seg[g] + len[g]) with
//! seg[g] = P[g] + ((row[g] - P[g]) & 3) and P[g] = sum over h < g of (len[h] + 3): the launch start keeps
//! the table row's alignment mod 4 (a lane's four rows stay 16-byte aligned), and the padding rows between
//! ranges are never live. Ranges past nq are empty at the end. Writes dims = {launch rows, tiles} and
//! clears the kept count.
__global__ void __launch_bounds__(kLayoutBlock) QueueLayoutKernel(const uint64_t *queue, uint32_t qmax,
                                                                uint64_t tile_rows, uint64_t *seg, uint64_t *seg_row,
                                                                uint64_t *seg_len, uint64_t *dims, uint64_t *count) {
	__shared__ uint64_t s_scan[kLayoutBlock / kWave + 1];
	const uint64_t nq = queue[0] < qmax ? queue[0] : qmax;
	const uint64_t *rows = queue + 1;
	const uint64_t *lens = queue + 1 + qmax;
	const uint32_t per = (qmax + kLayoutBlock - 1) / kLayoutBlock;
	const uint32_t g0 = threadIdx.x * per;
	uint64_t mine = 0;
	for (uint32_t i = 0; i < per; i++) {
		const uint32_t g = g0 + i;
		mine += g < nq ? lens[g] + 3 : 0;
	}
	uint64_t total = 0;
	uint64_t p = BlockExclusiveScan<uint64_t, kLayoutBlock>(mine, s_scan, &total);
	for (uint32_t i = 0; i < per; i++) {
		const uint32_t g = g0 + i;
		if (g >= qmax) {
			break;
		}
		if (g < nq) {
			seg[g] = p + ((rows[g] - p) & 3);
			seg_row[g] = rows[g];
			seg_len[g] = lens[g];
			p += lens[g] + 3;
		} else {
			seg[g] = total;
			seg_row[g] = 0;
			seg_len[g] = 0;
		}
	}
	if (threadIdx.x == 0) {
		seg[qmax] = total;
		dims[0] = total;
		dims[1] = (total + tile_rows - 1) / tile_rows;
		*count = 0;
	}
}

//! Zeroes the filter's look-back words and its ticket: state[0 .. dims[1]] (the ticket follows the tiles).
__global__ void ClearStateKernel(uint64_t *state, const uint64_t *dims) {
	const uint64_t words = dims[1] + 1;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
	     i += (uint64_t)gridDim.x * blockDim.x) {
		state[i] = 0;
	}
}

//! starts[g] = first output group of segment g (g = 0 .. nseg), by binary search over the groups' segment
//! column (ascending); the group count is read from device memory.
__global__ void SegmentStartsCountedKernel(const int32_t *seg_sorted, const uint64_t *groups, uint32_t nseg,
                                           uint64_t *starts) {
	const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g > nseg) {
		return;
	}
	uint64_t lo = 0, hi = *groups;
	while (lo < hi) {
		const uint64_t mid = (lo + hi) >> 1;
		if ((uint32_t)seg_sorted[mid] < g) {
			lo = mid + 1;
		} else {
			hi = mid;
		}
	}
	starts[g] = lo;
}

FusedExpr IdentityExpr(int slot, int vtype) {
	FusedExpr e;
	const std::string t = FusedCType(vtype);
	e.body = "((" + t + "*)out)[idx] = ((const " + t + "*)inputs[0])[row];";
	e.slots = {slot};
	e.out_type = vtype;
	return e;
}

} // namespace

//! Scratch that records the plan's allocations on its first pass (blocks from the device pool, held until
//! the plan is destroyed) and hands the same blocks back, in the same order, on the capture pass, which
//! therefore allocates nothing.
class BatchedGraphPlan::ReplayScratch : public Scratch {
public:
	explicit ReplayScratch(hipStream_t stream) : arena_(stream) {
	}
	void *Alloc(size_t bytes) override {
		if (!replay_) {
			void *p = arena_.Alloc(bytes);
			blocks_.push_back({p, bytes});
			return p;
		}
		if (next_ >= blocks_.size() || blocks_[next_].second < bytes) {
			throw std::runtime_error("batched plan: the capture pass asked for memory the sizing pass did not");
		}
		return blocks_[next_++].first;
	}
	void StartReplay() {
		replay_ = true;
		next_ = 0;
	}
	size_t Bytes() const {
		size_t b = 0;
		for (auto &kv : blocks_) {
			b += kv.second;
		}
		return b;
	}

private:
	PoolArena arena_;
	std::vector<std::pair<void *, size_t>> blocks_;
	bool replay_ = false;
	size_t next_ = 0;
};

BatchedGraphPlan::BatchedGraphPlan(const BatchedPlanArgs &args) : args_(args) {
	const auto &cols = args_.cols;
	if (args_.max_queries == 0 || args_.max_queries >= (1u << 30)) {
		throw std::runtime_error("batched plan: 1 .. 2^30 - 1 row ranges per batch");
	}
	if (args_.max_rows == 0 || args_.max_rows + 3ull * args_.max_queries >= (1ull << 30)) {
		throw std::runtime_error("batched plan: the row bound plus 3 rows per range must stay below 2^30");
	}
	if (args_.key_col < 0 || args_.key_col >= (int)cols.size() || !TypeIsInteger(cols[args_.key_col].type)) {
		throw std::runtime_error("batched plan: the group key must be an integer/boolean column");
	}
	if (args_.preds.empty()) {
		FusedExpr all; // ranges are compacted through the predicate path (padding rows are never kept)
		all.body = "((bool*)out)[idx] = true;";
		all.out_type = VT_BOOLEAN;
		args_.preds.push_back(all);
	}
	for (auto &c : cols) {
		if (c.rows != args_.table_rows) {
			throw std::runtime_error("batched plan: column row count mismatch");
		}
	}
	VGPU_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
	scratch_.reset(new ReplayScratch(stream_));
	try {
		Prepare();
	} catch (...) {
		Release();
		throw;
	}
}

BatchedGraphPlan::~BatchedGraphPlan() {
	Release();
}

void BatchedGraphPlan::Release() {
	if (exec_ != nullptr) {
		(void)hipGraphExecDestroy(exec_);
		exec_ = nullptr;
	}
	if (graph_ != nullptr) {
		(void)hipGraphDestroy(graph_);
		graph_ = nullptr;
	}
	if (stream_ != nullptr) {
		(void)hipStreamSynchronize(stream_);
	}
	scratch_.reset(); // pool blocks go back tagged with stream_
	if (queue_host_ != nullptr) {
		(void)hipHostFree(queue_host_);
		queue_host_ = nullptr;
	}
	if (result_host_ != nullptr) {
		(void)hipHostFree(result_host_);
		result_host_ = nullptr;
	}
	if (stream_ != nullptr) {
		DevicePool::Instance().ForgetStream(stream_);
		(void)hipStreamDestroy(stream_);
		stream_ = nullptr;
	}
}

void BatchedGraphPlan::Prepare() {
	const auto &cols = args_.cols;
	const uint32_t Q = args_.max_queries;
	// Generated filter: predicates, then segment id, key and value columns compacted (persistent form).
	FusedSpec spec;
	spec.fast_math = args_.fast_math;
	spec.ranges = true;
	spec.persistent = true;
	std::map<int, int> slot_of;
	auto slot = [&](int col) {
		if (col < 0 || col >= (int)cols.size()) {
			throw std::runtime_error("batched plan: column index out of range");
		}
		auto it = slot_of.find(col);
		if (it != slot_of.end()) {
			return it->second;
		}
		const int s = (int)spec.cols.size();
		FusedColumn fc;
		fc.vtype = cols[col].type;
		spec.cols.push_back(fc);
		col_ptrs_.push_back(cols[col].data);
		slot_of.emplace(col, s);
		return s;
	};
	for (auto p : args_.preds) {
		for (auto &s : p.slots) {
			s = slot(s);
		}
		spec.predicates.push_back(p);
	}
	const int seg_slot = (int)spec.cols.size();
	{
		FusedColumn fc;
		fc.vtype = VT_INT32;
		fc.segment_id = true;
		spec.cols.push_back(fc);
		col_ptrs_.push_back(nullptr);
	}
	spec.projections.push_back(IdentityExpr(seg_slot, VT_INT32));
	spec.projections.push_back(IdentityExpr(slot(args_.key_col), cols[args_.key_col].type));
	for (int v : args_.value_cols) {
		spec.projections.push_back(IdentityExpr(slot(v), cols[v].type));
	}
	if (spec.cols.size() > (size_t)kFusedMaxCols || spec.projections.size() > (size_t)kFusedMaxOuts) {
		throw std::runtime_error("batched plan: at most 16 columns");
	}
	kernel_ = GetFusedKernel(spec);

	// Composite key (segment, key) from the bounds: segments 0 .. Q-1, keys over the whole column.
	if (args_.table_rows > 0) {
		PoolArena tmp(stream_);
		const void *kc[1] = {cols[args_.key_col].data};
		int kt[1] = {cols[args_.key_col].type};
		uint64_t kmin[1], kmax[1];
		KeyRanges(kc, kt, 1, args_.table_rows, kmin, kmax, tmp, stream_);
		key_min_ = kmin[0];
		key_max_ = kmax[0];
	}
	std::vector<DevCol> keys(2);
	keys[0].type = VT_INT32;
	keys[1].type = cols[args_.key_col].type;
	const uint64_t mins[2] = {0ull ^ (1ull << 63), key_min_};
	const uint64_t maxs[2] = {(uint64_t)(Q - 1) ^ (1ull << 63), key_max_};
	group_bound_ = PlanGroupByRanges(keys, args_.max_rows, mins, maxs, pack_);
	if (group_bound_ > args_.capacity) {
		throw std::runtime_error("batched plan: output capacity " + std::to_string(args_.capacity) +
		                         " is below the group bound " + std::to_string(group_bound_));
	}

	// Host staging (pinned, so the graph's copy nodes read and write it directly).
	queue_words_ = 1 + 2ull * Q;
	result_words_ = Q + 2ull; // starts[0 .. Q], then the group count
	VGPU_HIP(hipHostMalloc(reinterpret_cast<void **>(&queue_host_), queue_words_ * 8, hipHostMallocDefault));
	VGPU_HIP(hipHostMalloc(reinterpret_cast<void **>(&result_host_), result_words_ * 8, hipHostMallocDefault));
	std::fill(queue_host_, queue_host_ + queue_words_, 0ull); // the sizing pass runs an empty batch
	std::fill(result_host_, result_host_ + result_words_, 0ull);

	// Pass 1 (uncaptured, empty batch): every allocation is made and recorded, every kernel loaded.
	Enqueue(stream_);
	VGPU_HIP(hipStreamSynchronize(stream_));
	footprint_ = scratch_->Bytes();
	// Pass 2: the same sequence captured; the scratch hands back pass 1's blocks.
	scratch_->StartReplay();
	VGPU_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
	try {
		Enqueue(stream_);
	} catch (...) {
		hipGraph_t partial = nullptr;
		(void)hipStreamEndCapture(stream_, &partial);
		if (partial != nullptr) {
			(void)hipGraphDestroy(partial);
		}
		throw;
	}
	VGPU_HIP(hipStreamEndCapture(stream_, &graph_));
	VGPU_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
	size_t nodes = 0;
	VGPU_HIP(hipGraphGetNodes(graph_, nullptr, &nodes));
	std::vector<hipGraphNode_t> list(nodes);
	VGPU_HIP(hipGraphGetNodes(graph_, list.data(), &nodes));
	for (auto n : list) {
		hipGraphNodeType t;
		VGPU_HIP(hipGraphNodeGetType(n, &t));
		if (t == hipGraphNodeTypeKernel) {
			kernel_nodes_++;
		} else if (t == hipGraphNodeTypeMemcpy) {
			copy_nodes_++;
		} else {
			other_nodes_++;
		}
	}
}

void BatchedGraphPlan::Enqueue(hipStream_t s) {
	const auto &cols = args_.cols;
	const uint32_t Q = args_.max_queries;
	Scratch &scratch = *scratch_;
	const uint64_t vbound = args_.max_rows + 3ull * Q; // launch rows: real rows plus at most 3 padding per range
	auto *queue = static_cast<uint64_t *>(scratch.Alloc(queue_words_ * 8));
	auto *layout = static_cast<uint64_t *>(scratch.Alloc((3ull * Q + 1 + 3) * 8));
	uint64_t *seg = layout;
	uint64_t *seg_row = layout + Q + 1;
	uint64_t *seg_len = layout + 2ull * Q + 1;
	uint64_t *dims = layout + 3ull * Q + 1; // {launch rows, tiles}, then the kept count
	uint64_t *kept = dims + 2;
	auto *result = static_cast<uint64_t *>(scratch.Alloc(result_words_ * 8));
	const size_t state_bytes = FusedKernel::StateBytes(vbound);
	auto *state = static_cast<uint64_t *>(scratch.Alloc(state_bytes));
	std::vector<void *> outs;
	outs.push_back(scratch.Alloc(args_.max_rows * 4));
	outs.push_back(scratch.Alloc(args_.max_rows * TypeWidth(cols[args_.key_col].type)));
	for (int v : args_.value_cols) {
		outs.push_back(scratch.Alloc(args_.max_rows * TypeWidth(cols[v].type)));
	}

	VGPU_HIP(hipMemcpyAsync(queue, queue_host_, queue_words_ * 8, hipMemcpyHostToDevice, s));
	hipLaunchKernelGGL(QueueLayoutKernel, dim3(1), dim3(kLayoutBlock), 0, s, queue, Q, kernel_->tile_rows(), seg,
	                   seg_row, seg_len, dims, kept);
	VGPU_HIP(hipGetLastError());
	{
		const uint64_t words = (vbound + kernel_->tile_rows() - 1) / kernel_->tile_rows() + 1;
		const unsigned grid = (unsigned)std::min<uint64_t>((words + 255) / 256, 1024);
		hipLaunchKernelGGL(ClearStateKernel, dim3(grid), dim3(256), 0, s, state, dims);
		VGPU_HIP(hipGetLastError());
	}
	FusedLaunch la;
	la.cols = col_ptrs_.data();
	la.outs = outs.data();
	la.n = vbound;
	la.capacity = args_.max_rows;
	la.count_device = kept;
	la.seg_offsets = seg;
	la.nseg = Q;
	la.seg_rows = seg_row;
	la.seg_lens = seg_len;
	la.dims_device = dims;
	kernel_->Launch(la, state, state_bytes, s);

	// GROUP BY (segment, key) over the kept rows, their count read on the device.
	std::vector<DevCol> keys(2), values;
	keys[0].data = outs[0];
	keys[0].type = VT_INT32;
	keys[0].rows = args_.max_rows;
	keys[1].data = outs[1];
	keys[1].type = cols[args_.key_col].type;
	keys[1].rows = args_.max_rows;
	for (size_t i = 0; i < args_.value_cols.size(); i++) {
		DevCol d;
		d.data = outs[2 + i];
		d.type = cols[args_.value_cols[i]].type;
		d.rows = args_.max_rows;
		values.push_back(d);
	}
	KeyPack pack = pack_;
	pack.col[0] = outs[0];
	pack.col[1] = outs[1];
	std::vector<void *> fixed = {args_.seg_out, args_.key_out};
	fixed.insert(fixed.end(), args_.agg_out.begin(), args_.agg_out.end());
	size_t next = 0;
	OutAlloc alloc = [&](size_t) { return fixed.at(next++); };
	std::vector<void *> key_out, agg_out;
	uint64_t *groups = result + Q + 1;
	GroupByDevice(keys, values, args_.aggs, args_.max_rows, pack, std::max<uint64_t>(args_.capacity, 1), alloc,
	              key_out, agg_out, TuneSort(pack.total_bits, args_.max_rows), scratch, s, false, groups, kept);
	hipLaunchKernelGGL(SegmentStartsCountedKernel, dim3((Q + 1 + 255) / 256), dim3(256), 0, s,
	                   static_cast<const int32_t *>(args_.seg_out), groups, Q, result);
	VGPU_HIP(hipGetLastError());
	VGPU_HIP(hipMemcpyAsync(result_host_, result, result_words_ * 8, hipMemcpyDeviceToHost, s));
}

uint64_t BatchedGraphPlan::Run(const uint64_t *row_starts, const uint64_t *row_lens, uint32_t nq,
                               uint64_t *seg_group_offsets, hipStream_t stream) {
	const uint32_t Q = args_.max_queries;
	if (nq > Q) {
		throw std::runtime_error("batched plan: " + std::to_string(nq) + " row ranges exceed the plan's " +
		                         std::to_string(Q));
	}
	uint64_t total = 0;
	for (uint32_t g = 0; g < nq; g++) {
		if (row_starts[g] > args_.table_rows || row_lens[g] > args_.table_rows - row_starts[g]) {
			throw std::runtime_error("batched plan: range " + std::to_string(g) + " [" + std::to_string(row_starts[g]) +
			                         ", +" + std::to_string(row_lens[g]) + ") is past the table's " +
			                         std::to_string(args_.table_rows) + " rows");
		}
		total += row_lens[g];
		if (total > args_.max_rows) {
			throw std::runtime_error("batched plan: the batch's rows exceed the plan's bound of " +
			                         std::to_string(args_.max_rows));
		}
	}
	// The previous run waited for its graph, so the staging buffers are free.
	queue_host_[0] = nq;
	std::copy(row_starts, row_starts + nq, queue_host_ + 1);
	std::fill(queue_host_ + 1 + nq, queue_host_ + 1 + Q, 0ull);
	std::copy(row_lens, row_lens + nq, queue_host_ + 1 + Q);
	std::fill(queue_host_ + 1 + Q + nq, queue_host_ + 1 + 2ull * Q, 0ull);
	VGPU_HIP(hipGraphLaunch(exec_, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	const uint64_t groups = result_host_[Q + 1];
	if (seg_group_offsets != nullptr) {
		std::copy(result_host_, result_host_ + nq + 1, seg_group_offsets);
	}
	return groups;
}

} // namespace vgpu
