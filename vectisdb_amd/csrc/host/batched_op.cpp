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
#include "device_pool.h"
#include "operators.h"
#include "../hip/join_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
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
	const bool ranged = args.range_rows_host != nullptr;
	if (ranged != (args.range_lens_host != nullptr)) {
		throw std::runtime_error("batched group by: row ranges need both starts and lengths");
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
	if (ranged && spec.predicates.empty()) {
		// row ranges are compacted through the predicate path (padding rows are never kept)
		FusedExpr all;
		all.body = "((bool*)out)[idx] = true;";
		all.out_type = VT_BOOLEAN;
		spec.predicates.push_back(all);
	}
	spec.ranges = ranged;
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
	// The composite key's ranges come out of the same pass (no separate read of the compacted keys).
	spec.range_slots = {seg_slot, slot(args.key_col)};
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

	DeviceRanges ranges;
	if (ranged) {
		ranges = UploadRanges(args.range_rows_host, args.range_lens_host, args.nseg, rows, scratch, stream);
	}
	const uint64_t n = ranged ? ranges.nvirt : rows; // launch rows
	std::vector<void *> outs;
	outs.push_back(scratch.Alloc(std::max<uint64_t>(n, 1) * 4));
	outs.push_back(scratch.Alloc(std::max<uint64_t>(n, 1) * TypeWidth(cols[args.key_col].type)));
	for (int v : args.value_cols) {
		outs.push_back(scratch.Alloc(std::max<uint64_t>(n, 1) * TypeWidth(cols[v].type)));
	}
	// count, then {min, ~max} of the segment id and of the key over the kept rows
	auto *count = static_cast<uint64_t *>(scratch.Alloc(5 * 8));
	uint64_t kept = n;
	FusedLaunch la;
	la.range_device = count + 1;
	la.cols = ptrs.data();
	la.outs = outs.data();
	la.n = n;
	la.seg_offsets = ranged ? ranges.seg : args.seg_offsets_dev;
	la.nseg = args.nseg;
	la.seg_rows = ranges.rows;
	la.seg_lens = ranges.lens;
	void *state = nullptr;
	size_t state_bytes = 0;
	if (!spec.predicates.empty()) {
		la.count_device = count;
		state_bytes = FusedKernel::StateBytes(n);
		state = scratch.Alloc(state_bytes);
	}
	kernel->Launch(la, state, state_bytes, stream);
	uint64_t words[5] = {};
	if (spec.predicates.empty()) {
		VGPU_HIP(hipMemcpyAsync(words + 1, count + 1, 4 * 8, hipMemcpyDeviceToHost, stream));
	} else {
		VGPU_HIP(hipMemcpyAsync(words, count, 5 * 8, hipMemcpyDeviceToHost, stream));
	}
	VGPU_HIP(hipStreamSynchronize(stream));
	if (!spec.predicates.empty()) {
		kept = n == 0 ? 0 : words[0];
	}

	if (kept == 0) { // nothing survived: every query has zero groups
		if (args.seg_group_offsets_host != nullptr) {
			std::fill(args.seg_group_offsets_host, args.seg_group_offsets_host + args.nseg + 1, 0ull);
		}
		return 0;
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
	const uint64_t mins[2] = {words[1], words[3]}, maxs[2] = {~words[2], ~words[4]};
	const uint64_t bound = PlanGroupByRanges(keys, kept, mins, maxs, pack);
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

DeviceRanges UploadRanges(const uint64_t *rows, const uint64_t *lens, uint32_t nseg, uint64_t table_rows,
                          Scratch &scratch, hipStream_t stream) {
	if (nseg == 0) {
		throw std::runtime_error("row ranges: at least one range");
	}
	std::vector<uint64_t> host(3ull * nseg + 1);
	uint64_t *seg = host.data();
	uint64_t v = 0;
	for (uint32_t g = 0; g < nseg; g++) {
		if (rows[g] > table_rows || lens[g] > table_rows - rows[g]) {
			throw std::runtime_error("row ranges: range " + std::to_string(g) + " [" + std::to_string(rows[g]) +
			                         ", +" + std::to_string(lens[g]) + ") is past the table's " +
			                         std::to_string(table_rows) + " rows");
		}
		v += (rows[g] - v) & 3; // launch start = table start (mod 4)
		seg[g] = v;
		v += lens[g];
	}
	seg[nseg] = v;
	std::copy(rows, rows + nseg, host.begin() + nseg + 1);
	std::copy(lens, lens + nseg, host.begin() + 2 * nseg + 1);
	auto *dev = static_cast<uint64_t *>(scratch.Alloc(host.size() * 8));
	VGPU_HIP(hipMemcpyAsync(dev, host.data(), host.size() * 8, hipMemcpyHostToDevice, stream));
	VGPU_HIP(hipStreamSynchronize(stream)); // `host` is pageable and local
	DeviceRanges r;
	r.seg = dev;
	r.rows = dev + nseg + 1;
	r.lens = dev + 2 * nseg + 1;
	r.nvirt = v;
	r.nseg = nseg;
	return r;
}

uint64_t BatchedFilterProjectDevice(const std::vector<DevCol> &cols, const DeviceRanges &ranges,
                                    const std::vector<FusedExpr> &preds, const std::vector<FusedExpr> &projs,
                                    bool fast_math, std::vector<void *> &outs, std::vector<uint64_t> &kept_offsets,
                                    Scratch &out_scratch, Scratch &scratch, hipStream_t stream) {
	FusedSpec spec;
	spec.fast_math = fast_math;
	spec.ranges = true;
	std::map<int, int> slot_of;
	std::vector<const void *> ptrs;
	auto slot = [&](int col) {
		if (col < 0 || col >= (int)cols.size()) {
			throw std::runtime_error("batched filter/project: expression input out of range");
		}
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
	for (auto p : preds) {
		for (auto &s : p.slots) {
			s = slot(s);
		}
		spec.predicates.push_back(p);
	}
	if (spec.predicates.empty()) {
		FusedExpr all;
		all.body = "((bool*)out)[idx] = true;";
		all.out_type = VT_BOOLEAN;
		spec.predicates.push_back(all);
	}
	for (auto p : projs) {
		for (auto &s : p.slots) {
			s = slot(s);
		}
		spec.projections.push_back(p);
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
	if (spec.cols.size() > (size_t)kFusedMaxCols || spec.projections.size() > (size_t)kFusedMaxOuts) {
		throw std::runtime_error("batched filter/project: at most 15 columns and 15 projections");
	}
	auto kernel = GetFusedKernel(spec);
	const uint64_t n = ranges.nvirt;
	outs.clear();
	for (auto &p : projs) {
		outs.push_back(out_scratch.Alloc(std::max<uint64_t>(n, 1) * TypeWidth(p.out_type)));
	}
	auto *seg_out = static_cast<int32_t *>(scratch.Alloc(std::max<uint64_t>(n, 1) * 4));
	std::vector<void *> all_outs = outs;
	all_outs.push_back(seg_out);
	auto *count = static_cast<uint64_t *>(scratch.Alloc(8));
	const size_t state_bytes = FusedKernel::StateBytes(n);
	void *state = scratch.Alloc(std::max<size_t>(state_bytes, 8));
	FusedLaunch la;
	la.cols = ptrs.data();
	la.outs = all_outs.data();
	la.n = n;
	la.count_device = count;
	la.seg_offsets = ranges.seg;
	la.nseg = ranges.nseg;
	la.seg_rows = ranges.rows;
	la.seg_lens = ranges.lens;
	kernel->Launch(la, state, state_bytes, stream);
	auto *starts = static_cast<uint64_t *>(scratch.Alloc((ranges.nseg + 1) * 8));
	uint64_t kept = 0;
	VGPU_HIP(hipMemcpyAsync(&kept, count, 8, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	if (n == 0) {
		kept = 0;
	}
	hipLaunchKernelGGL(SegmentStartsKernel, dim3((ranges.nseg + 1 + 255) / 256), dim3(256), 0, stream, seg_out, kept,
	                   ranges.nseg, starts);
	VGPU_HIP(hipGetLastError());
	kept_offsets.assign(ranges.nseg + 1, 0);
	VGPU_HIP(hipMemcpyAsync(kept_offsets.data(), starts, (ranges.nseg + 1) * 8, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	return kept;
}

JoinIndex::~JoinIndex() {
	auto &pool = DevicePool::Instance();
	for (void *p : {keys, static_cast<void *>(count), static_cast<void *>(sum), static_cast<void *>(dense)}) {
		if (p != nullptr) {
			pool.FreeAny(p);
		}
	}
}

std::unique_ptr<JoinIndex> BuildJoinIndex(const DevCol &key, const DevCol &payload, hipStream_t stream,
                                          bool dense_only) {
	const bool has_payload = payload.data != nullptr;
	if (!TypeIsInteger(key.type) || key.type == VT_BOOLEAN || (has_payload && !TypeIsInteger(payload.type))) {
		throw std::runtime_error("join index: integer key (INT16/32/64) and integer payload");
	}
	if (has_payload && key.rows != payload.rows) {
		throw std::runtime_error("join index: key and payload of equal length");
	}
	auto index = std::unique_ptr<JoinIndex>(new JoinIndex());
	index->key_type = key.type;
	index->build_rows = key.rows;
	if (key.rows == 0) {
		return index;
	}
	PoolArena arena(stream);
	std::vector<DevCol> keys = {key};
	KeyPack pack;
	uint64_t span = 0;
	const uint64_t bound = PlanGroupBy(keys, key.rows, pack, arena, stream, nullptr, &span);
	if (dense_only && (span == 0 || span > 2 * key.rows + (1ull << 20))) {
		return index; // too sparse for the direct-address table: not built
	}
	auto &pool = DevicePool::Instance();
	const uint64_t cap = std::max<uint64_t>(bound, 1);
	index->keys = pool.Alloc(cap * TypeWidth(key.type), stream);
	index->count = static_cast<uint64_t *>(pool.Alloc(cap * 8, stream));
	index->sum = static_cast<uint64_t *>(pool.Alloc(cap * 16, stream));
	std::vector<void *> fixed = {index->keys, index->count, index->sum};
	size_t next = 0;
	OutAlloc alloc = [&](size_t) { return fixed.at(next++); };
	std::vector<AggReq> aggs = {{AGG_COUNT_STAR, -1, VT_INT64}};
	std::vector<DevCol> values;
	if (has_payload) {
		aggs.push_back({AGG_SUM, 0, VT_HUGEINT});
		values.push_back(payload);
	} else {
		FillU64(index->sum, 0ull, cap * 2, stream); // no payload: every sum is 0
	}
	std::vector<void *> ko, ao;
	SortTuning tuning = TuneSort(pack.total_bits, key.rows);
	index->distinct =
	    GroupByDevice(keys, values, aggs, key.rows, pack, cap, alloc, ko, ao, tuning, arena, stream, true);
	if (index->distinct > 0) {
		// keys ascend: the span is last - first, as sign-extended 64-bit values
		const int w = TypeWidth(key.type);
		int64_t first = 0, last = 0;
		uint8_t buf[2][8] = {};
		VGPU_HIP(hipMemcpyAsync(buf[0], index->keys, w, hipMemcpyDeviceToHost, stream));
		VGPU_HIP(hipMemcpyAsync(buf[1], static_cast<const uint8_t *>(index->keys) + (index->distinct - 1) * w, w,
		                        hipMemcpyDeviceToHost, stream));
		VGPU_HIP(hipStreamSynchronize(stream));
		for (int i = 0; i < 2; i++) {
			int64_t v = 0;
			if (w == 2) {
				int16_t t;
				std::memcpy(&t, buf[i], 2);
				v = t;
			} else if (w == 4) {
				int32_t t;
				std::memcpy(&t, buf[i], 4);
				v = t;
			} else {
				std::memcpy(&v, buf[i], 8);
			}
			(i == 0 ? first : last) = v;
		}
		const uint64_t span = (uint64_t)last - (uint64_t)first + 1; // exact: last >= first
		if (span != 0 && span <= 2 * index->distinct + (1ull << 20)) {
			auto *dense = static_cast<uint64_t *>(pool.Alloc(span * 8, stream));
			auto *flag = static_cast<unsigned int *>(arena.Alloc(4));
			if (BuildDenseIndex(key.type, index->keys, index->count, index->sum, index->distinct, (uint64_t)first,
			                    span, dense, flag, stream)) {
				index->dense = dense;
				index->dense_min = (uint64_t)first;
				index->dense_span = span;
			} else {
				pool.FreeAny(dense);
			}
		}
	}
	return index;
}

std::vector<JoinAggValue> BatchedJoinAggDevice(const DevCol &probe_key, const DeviceRanges &ranges,
                                               const JoinIndex &index, Scratch &scratch, hipStream_t stream) {
	if (probe_key.type != index.key_type) {
		throw std::runtime_error("batched join: probe and build keys must have the same type");
	}
	auto *acc = static_cast<uint64_t *>(scratch.Alloc(3ull * ranges.nseg * 8));
	RangeProbeAgg(probe_key.type, probe_key.data, ranges.seg, ranges.rows, ranges.lens, ranges.nseg, ranges.nvirt,
	              index.keys, index.count, index.sum, index.distinct, index.dense, index.dense_min, index.dense_span,
	              acc, stream);
	std::vector<uint64_t> host(3ull * ranges.nseg);
	VGPU_HIP(hipMemcpyAsync(host.data(), acc, host.size() * 8, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	std::vector<JoinAggValue> out(2ull * ranges.nseg);
	for (uint32_t g = 0; g < ranges.nseg; g++) {
		out[2 * g].lo = host[3 * g + 1];
		out[2 * g].hi = host[3 * g + 2];
		out[2 * g].kind = JA_SUM_I128;
		out[2 * g + 1].lo = host[3 * g];
		out[2 * g + 1].kind = JA_COUNT;
	}
	return out;
}

} // namespace vgpu
