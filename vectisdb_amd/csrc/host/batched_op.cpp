// batched_op.cpp -- the batched form: many filter + GROUP BY queries, each over its own row group of one
// device-resident table, executed as ONE pipeline instead of one pipeline per query.
//
// The row groups are segments of the table (segment g covers rows [offsets[g], offsets[g+1])). One
// generated kernel evaluates every query's predicates over the whole table and writes the surviving rows
// with their segment index as an extra INT32 column (the segment comes from the offsets, not from memory);
// one GROUP BY over the composite key (segment, key) then sorts and reduces every query's groups in the
// same radix passes and segmented reduction. Tiles are handed out by atomic tickets, so row groups of very
// different sizes load-balance across the GPU the way one large input does. Each query's groups are the
// contiguous run of output rows with its segment index (group order within a query = ascending key), the
// same rows the query produces alone.
#include "../engine/fused.h"
#include "operators.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>

namespace vgpu {

namespace {

const char *IdentityType(int vtype) {
	switch (vtype) {
	case VT_INT16:
		return "short";
	case VT_INT32:
		return "int";
	case VT_INT64:
		return "long long";
	case VT_FLOAT32:
		return "float";
	case VT_FLOAT64:
		return "double";
	case VT_BOOLEAN:
		return "bool";
	default:
		throw std::runtime_error("batched group by: column type cannot be carried by a generated kernel");
	}
}

FusedExpr Identity(int slot, int vtype) {
	FusedExpr e;
	const std::string t = IdentityType(vtype);
	e.body = "((" + t + "*)out)[idx] = ((const " + t + "*)inputs[0])[row];";
	e.slots = {slot};
	e.out_type = vtype;
	return e;
}

__global__ void SegmentStartsKernel(const int32_t *seg_sorted, uint64_t n, uint32_t nseg, uint64_t *starts) {
	const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g > nseg) {
		return;
	}
	// first index whose segment is >= g
	uint64_t lo = 0, hi = n;
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

} // namespace

uint64_t BatchedFilterGroupByDevice(const BatchedGroupByArgs &args, Scratch &scratch, hipStream_t stream) {
	const auto &cols = args.cols;
	const uint64_t rows = args.rows;
	if (args.nseg == 0 || args.nseg >= (1u << 30)) {
		throw std::runtime_error("batched group by: 1 .. 2^30 - 1 row groups");
	}
	if (args.key_col < 0 || args.key_col >= (int)cols.size() || !TypeIsInteger(cols[args.key_col].type)) {
		throw std::runtime_error("batched group by: the group key must be an integer/boolean column");
	}
	for (auto &c : cols) {
		if (c.rows != rows) {
			throw std::runtime_error("batched group by: column row count mismatch");
		}
	}
	// Generated kernel: predicates over the table, then segment id, key and each value column compacted.
	FusedSpec spec;
	spec.fast_math = args.fast_math;
	std::map<int, int> slot_of;
	std::vector<const void *> ptrs;
	auto slot = [&](int col) {
		auto it = slot_of.find(col);
		if (it != slot_of.end()) {
			return it->second;
		}
		const int s = (int)spec.cols.size();
		FusedColumn fc;
		fc.vtype = cols[col].type;
		spec.cols.push_back(fc);
		ptrs.push_back(cols[col].data);
		slot_of.emplace(col, s);
		return s;
	};
	for (auto p : args.preds) {
		for (auto &s : p.slots) {
			if (s < 0 || s >= (int)cols.size()) {
				throw std::runtime_error("batched group by: predicate input out of range");
			}
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
		ptrs.push_back(nullptr);
	}
	spec.projections.push_back(Identity(seg_slot, VT_INT32));
	spec.projections.push_back(Identity(slot(args.key_col), cols[args.key_col].type));
	for (int v : args.value_cols) {
		if (v < 0 || v >= (int)cols.size()) {
			throw std::runtime_error("batched group by: value column out of range");
		}
		spec.projections.push_back(Identity(slot(v), cols[v].type));
	}
	if (spec.cols.size() > (size_t)kFusedMaxCols || spec.projections.size() > (size_t)kFusedMaxOuts) {
		throw std::runtime_error("batched group by: at most 16 columns");
	}
	auto kernel = GetFusedKernel(spec);

	std::vector<void *> outs;
	outs.push_back(scratch.Alloc(std::max<uint64_t>(rows, 1) * 4));
	outs.push_back(scratch.Alloc(std::max<uint64_t>(rows, 1) * TypeWidth(cols[args.key_col].type)));
	for (int v : args.value_cols) {
		outs.push_back(scratch.Alloc(std::max<uint64_t>(rows, 1) * TypeWidth(cols[v].type)));
	}
	auto *count = static_cast<uint64_t *>(scratch.Alloc(8));
	uint64_t kept = rows;
	FusedLaunch la;
	la.cols = ptrs.data();
	la.outs = outs.data();
	la.n = rows;
	la.seg_offsets = args.seg_offsets_dev;
	la.nseg = args.nseg;
	void *state = nullptr;
	size_t state_bytes = 0;
	if (!spec.predicates.empty()) {
		la.count_device = count;
		state_bytes = FusedKernel::StateBytes(rows);
		state = scratch.Alloc(state_bytes);
	}
	kernel->Launch(la, state, state_bytes, stream);
	if (!spec.predicates.empty()) {
		VGPU_HIP(hipMemcpyAsync(&kept, count, 8, hipMemcpyDeviceToHost, stream));
		VGPU_HIP(hipStreamSynchronize(stream));
		if (rows == 0) {
			kept = 0;
		}
	}

	// GROUP BY (segment, key) over the surviving rows.
	std::vector<DevCol> keys(2), values;
	keys[0].data = outs[0];
	keys[0].type = VT_INT32;
	keys[0].rows = kept;
	keys[1].data = outs[1];
	keys[1].type = cols[args.key_col].type;
	keys[1].rows = kept;
	for (size_t i = 0; i < args.value_cols.size(); i++) {
		DevCol d;
		d.data = outs[2 + i];
		d.type = cols[args.value_cols[i]].type;
		d.rows = kept;
		values.push_back(d);
	}
	KeyPack pack;
	const uint64_t bound = PlanGroupBy(keys, kept, pack, scratch, stream);
	if (bound > args.capacity) {
		throw std::runtime_error("batched group by: output capacity " + std::to_string(args.capacity) +
		                         " is below the group bound " + std::to_string(bound));
	}
	std::vector<void *> key_out, agg_out;
	size_t next = 0;
	std::vector<void *> fixed = {args.seg_out, args.key_out};
	fixed.insert(fixed.end(), args.agg_out.begin(), args.agg_out.end());
	OutAlloc alloc = [&](size_t) { return fixed.at(next++); };
	SortTuning tuning = TuneSort(pack.total_bits, kept);
	const uint64_t groups = GroupByDevice(keys, values, args.aggs, kept, pack, std::max<uint64_t>(args.capacity, 1),
	                                      alloc, key_out, agg_out, tuning, scratch, stream, true);
	// Per-segment start of its groups (segments are ascending in the output).
	auto *starts = static_cast<uint64_t *>(scratch.Alloc((args.nseg + 1) * 8));
	hipLaunchKernelGGL(SegmentStartsKernel, dim3((args.nseg + 1 + 255) / 256), dim3(256), 0, stream,
	                   static_cast<const int32_t *>(args.seg_out), groups, args.nseg, starts);
	VGPU_HIP(hipGetLastError());
	if (args.seg_group_offsets_host != nullptr) {
		VGPU_HIP(hipMemcpyAsync(args.seg_group_offsets_host, starts, (args.nseg + 1) * 8, hipMemcpyDeviceToHost,
		                        stream));
	}
	VGPU_HIP(hipStreamSynchronize(stream));
	return groups;
}

uint64_t PerQueryFilterGroupByDevice(const BatchedGroupByArgs &args, const uint64_t *seg_offsets_host,
                                     Scratch &scratch, hipStream_t stream) {
	const auto &cols = args.cols;
	if (args.key_col < 0 || args.key_col >= (int)cols.size() || !TypeIsInteger(cols[args.key_col].type)) {
		throw std::runtime_error("per-query group by: the group key must be an integer/boolean column");
	}
	// One generated kernel (predicates, then key and values compacted), reused by every query.
	FusedSpec spec;
	spec.fast_math = args.fast_math;
	std::map<int, int> slot_of;
	std::vector<int> slot_col;
	auto slot = [&](int col) {
		auto it = slot_of.find(col);
		if (it != slot_of.end()) {
			return it->second;
		}
		const int s = (int)spec.cols.size();
		FusedColumn fc;
		fc.vtype = cols[col].type;
		spec.cols.push_back(fc);
		slot_col.push_back(col);
		slot_of.emplace(col, s);
		return s;
	};
	for (auto p : args.preds) {
		for (auto &s : p.slots) {
			s = slot(s);
		}
		spec.predicates.push_back(p);
	}
	spec.projections.push_back(Identity(slot(args.key_col), cols[args.key_col].type));
	for (int v : args.value_cols) {
		spec.projections.push_back(Identity(slot(v), cols[v].type));
	}
	auto kernel = GetFusedKernel(spec);
	uint64_t max_rows = 0;
	for (uint32_t q = 0; q < args.nseg; q++) {
		max_rows = std::max<uint64_t>(max_rows, seg_offsets_host[q + 1] - seg_offsets_host[q]);
	}
	std::vector<void *> outs;
	outs.push_back(scratch.Alloc(std::max<uint64_t>(max_rows, 1) * TypeWidth(cols[args.key_col].type)));
	for (int v : args.value_cols) {
		outs.push_back(scratch.Alloc(std::max<uint64_t>(max_rows, 1) * TypeWidth(cols[v].type)));
	}
	auto *count = static_cast<uint64_t *>(scratch.Alloc(8));
	void *state = scratch.Alloc(FusedKernel::StateBytes(std::max<uint64_t>(max_rows, 1)));
	uint64_t total = 0;
	for (uint32_t q = 0; q < args.nseg; q++) {
		if (args.seg_group_offsets_host != nullptr) {
			args.seg_group_offsets_host[q] = total;
		}
		const uint64_t r0 = seg_offsets_host[q];
		const uint64_t n = seg_offsets_host[q + 1] - r0;
		if (n == 0) {
			continue;
		}
		std::vector<const void *> ptrs;
		for (int c : slot_col) {
			ptrs.push_back(static_cast<const uint8_t *>(cols[c].data) + r0 * TypeWidth(cols[c].type));
		}
		FusedLaunch la;
		la.cols = ptrs.data();
		la.outs = outs.data();
		la.n = n;
		la.count_device = spec.predicates.empty() ? nullptr : count;
		kernel->Launch(la, state, FusedKernel::StateBytes(n), stream);
		uint64_t kept = n;
		if (!spec.predicates.empty()) {
			VGPU_HIP(hipMemcpyAsync(&kept, count, 8, hipMemcpyDeviceToHost, stream));
			VGPU_HIP(hipStreamSynchronize(stream));
		}
		if (kept == 0) {
			continue;
		}
		std::vector<DevCol> keys(1), values;
		keys[0].data = outs[0];
		keys[0].type = cols[args.key_col].type;
		keys[0].rows = kept;
		for (size_t i = 0; i < args.value_cols.size(); i++) {
			DevCol d;
			d.data = outs[1 + i];
			d.type = cols[args.value_cols[i]].type;
			d.rows = kept;
			values.push_back(d);
		}
		KeyPack pack;
		const uint32_t *hist8 = nullptr;
		const uint64_t bound = PlanGroupBy(keys, kept, pack, scratch, stream, &hist8);
		if (total + bound > args.capacity) {
			throw std::runtime_error("per-query group by: output capacity exceeded");
		}
		std::vector<void *> fixed = {static_cast<uint8_t *>(args.key_out) + total * TypeWidth(keys[0].type)};
		for (size_t i = 0; i < args.aggs.size(); i++) {
			fixed.push_back(static_cast<uint8_t *>(args.agg_out[i]) + total * TypeWidth(args.aggs[i].out_type));
		}
		size_t next = 0;
		OutAlloc alloc = [&](size_t) { return fixed.at(next++); };
		std::vector<void *> ko, ao;
		SortTuning tuning = TuneSort(pack.total_bits, kept);
		tuning.hist8 = hist8;
		const uint64_t g = GroupByDevice(keys, values, args.aggs, kept, pack, std::max<uint64_t>(bound, 1), alloc, ko,
		                                 ao, tuning, scratch, stream, true);
		FillU32(static_cast<uint32_t *>(args.seg_out) + total, q, g, stream);
		total += g;
	}
	if (args.seg_group_offsets_host != nullptr) {
		args.seg_group_offsets_host[args.nseg] = total;
	}
	VGPU_HIP(hipStreamSynchronize(stream));
	return total;
}

} // namespace vgpu
