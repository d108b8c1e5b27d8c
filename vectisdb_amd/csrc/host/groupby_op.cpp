// groupby_op.cpp -- host side of the sort-based GROUP BY: validation, key packing, component planning,
// then radix sort (keys + aggregate inputs as payload) and the one-pass segmented reduction.
//
// Semantics follow cuda_engine/src/operators/gpu_groupby.cu:224-274 (ValidateAggregate) and :303-592
// (ExecuteGroupByNode): integer/boolean keys only, output = group keys then aggregates, COUNT(*) and
// COUNT(col) INT64, MIN/MAX keep the input type, AVG DOUBLE. Extensions: SUM over integer columns
// produces HUGEINT (what DuckDB returns and what the reference refuses at :242-247), and MIN/MAX over
// FLOAT/DOUBLE order NaN above every value as DuckDB does.
#include "operators.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <utility>

namespace vgpu {

namespace {
bool IsFloat(int t) {
	return t == VT_FLOAT32 || t == VT_FLOAT64;
}

//! INT64 columns (address, rows) whose values left the INT32 range when narrowed: later GROUP BYs over them
//! carry them whole. A short FIFO: a recycled address costs at most one narrowed attempt more.
std::mutex g_wide_mu;
std::deque<std::pair<const void *, uint64_t>> g_wide;
constexpr size_t kWideMemory = 64;

bool KnownWide(const void *col, uint64_t rows) {
	std::lock_guard<std::mutex> lock(g_wide_mu);
	return std::find(g_wide.begin(), g_wide.end(), std::make_pair(col, rows)) != g_wide.end();
}

void NoteWide(const void *col, uint64_t rows) {
	std::lock_guard<std::mutex> lock(g_wide_mu);
	g_wide.emplace_back(col, rows);
	if (g_wide.size() > kWideMemory) {
		g_wide.pop_front();
	}
}
bool IsNumeric(int t) {
	return t == VT_INT16 || t == VT_INT32 || t == VT_INT64 || IsFloat(t);
}

std::atomic<int> g_run_mode {[] {
	const char *e = std::getenv("VGPU_RUN_REDUCE");
	return e == nullptr ? 1 : std::max(0, std::min(2, std::atoi(e)));
}()};

//! The last radix pass fused with the reduction pays when the partial states it may write (one per distinct
//! low-bits value and tile boundary, per value of the top digit), written once and read back by the record
//! combine, are fewer bytes than what the plain path moves after the last ranking: the rows the last pass
//! scatters, the keys the head count reads and the rows the segmented reduction reads -- and when there are at
//! least 8 rows per partial state: with fewer, each tile writes about one state into each of up to 256 digit
//! regions, partial-line writes that cost more than their bytes (the count-only side of the sort-merge join,
//! 2^28 rows over 2^24 keys: 2.0 ms fused against 1.8 ms plain).
bool RunReduceWorthIt(const KeyPack &pack, const ReduceSpec &rs, const std::vector<int> &widths, uint64_t rows) {
	const int mode = g_run_mode.load();
	const int B = std::max(pack.total_bits, 1);
	const int P = (B + 7) / 8;
	if (mode == 0 || P < 2 || rows == 0 || !RunReduceSupported(rs, (int)widths.size(), B > 32)) {
		return false;
	}
	if (mode == 2) {
		return true;
	}
	const int last_bits = B / P;
	const int shift = B - last_bits;
	const uint64_t max_packed = pack.max_packed != 0 ? pack.max_packed : (B >= 64 ? ~0ull : (1ull << B) - 1);
	const uint64_t records = RunReduceRecordBound(rows, shift, last_bits, max_packed);
	const double kw = B > 32 ? 8.0 : 4.0;
	double row_bytes = kw;
	for (int w : widths) {
		row_bytes += w;
	}
	return records * 8 <= rows && (double)records * (3.0 * kw + 32.0 * rs.ncomp) <= (double)rows * (2.0 * row_bytes + kw);
}
} // namespace

void SetRunReduceMode(int mode) {
	g_run_mode.store(std::max(0, std::min(2, mode)));
}

int RunReduceMode() {
	return g_run_mode.load();
}

void ValidateAggregateReq(const AggReq &agg, const std::vector<DevCol> &values) {
	if (agg.kind == AGG_COUNT_STAR) {
		if (agg.out_type != VT_INT64) {
			throw std::runtime_error("gpu_groupby: COUNT(*) must produce INT64");
		}
		return;
	}
	if (agg.input < 0 || agg.input >= (int)values.size()) {
		throw std::runtime_error("gpu_groupby: aggregate input column index out of range");
	}
	const int in = values[agg.input].type;
	switch (agg.kind) {
	case AGG_COUNT:
		if (agg.out_type != VT_INT64) {
			throw std::runtime_error("gpu_groupby: COUNT must produce INT64");
		}
		return;
	case AGG_SUM:
		if (IsFloat(in)) {
			if (agg.out_type != VT_FLOAT64) {
				throw std::runtime_error("gpu_groupby: SUM over FLOAT/DOUBLE must produce DOUBLE");
			}
			return;
		}
		if (in == VT_INT16 || in == VT_INT32 || in == VT_INT64) {
			if (agg.out_type != VT_HUGEINT && agg.out_type != VT_INT64 && agg.out_type != VT_FLOAT64) {
				throw std::runtime_error("gpu_groupby: SUM over an integer column must produce HUGEINT "
				                         "(DuckDB's result type), INT64 or DOUBLE");
			}
			return;
		}
		throw std::runtime_error("gpu_groupby: SUM supports numeric columns only");
	case AGG_AVG:
		if (agg.out_type != VT_FLOAT64) {
			throw std::runtime_error("gpu_groupby: AVG must produce DOUBLE");
		}
		if (!IsNumeric(in)) {
			throw std::runtime_error("gpu_groupby: AVG supports numeric columns only");
		}
		return;
	case AGG_MIN:
	case AGG_MAX:
		if (agg.out_type != in) {
			throw std::runtime_error("gpu_groupby: MIN/MAX output type must match its input column type");
		}
		if (!(TypeIsInteger(in) || IsFloat(in))) {
			throw std::runtime_error("gpu_groupby: MIN/MAX supports integer, boolean and floating columns");
		}
		return;
	default:
		throw std::runtime_error("gpu_groupby: unknown aggregate kind");
	}
}

uint64_t PlanGroupBy(const std::vector<DevCol> &keys, uint64_t rows, KeyPack &pack, Scratch &scratch,
                     hipStream_t stream, const uint32_t **hist8, uint64_t *span_out) {
	pack = KeyPack {};
	if (span_out != nullptr) {
		*span_out = 0;
	}
	if (hist8 != nullptr) {
		*hist8 = nullptr;
	}
	if (keys.size() == 1 && rows > 0 && TypeIsInteger(keys[0].type) && TuneSort(64, rows).radix_bits == 8) {
		// One read: min/max and every byte-digit histogram.
		auto *h = static_cast<uint32_t *>(scratch.Alloc((8 * 256 + 4) * sizeof(uint32_t)));
		uint64_t omin = 0, omax = 0;
		KeyRangeHist8(keys[0].data, keys[0].type, rows, &omin, &omax, h, stream);
		PlanKeyPackAligned(keys[0].data, keys[0].type, omin, omax, pack);
		if (span_out != nullptr) {
			*span_out = omax - omin + 1;
		}
		if (hist8 != nullptr) {
			*hist8 = h;
		}
		uint64_t range = omax - omin;
		return range < rows ? range + 1 : rows;
	}
	if (keys.empty()) {
		return 1;
	}
	if (keys.size() > (size_t)kMaxKeyCols) {
		throw std::runtime_error("gpu_groupby: at most 4 group key columns are supported");
	}
	std::vector<const void *> cols;
	std::vector<int> types;
	for (auto &k : keys) {
		if (!TypeIsInteger(k.type)) {
			throw std::runtime_error("gpu_groupby: group key is not an integer/boolean column, which this GPU "
			                         "path requires");
		}
		cols.push_back(k.data);
		types.push_back(k.type);
	}
	uint64_t mins[kMaxKeyCols] = {}, maxs[kMaxKeyCols] = {};
	if (rows == 0) {
		PlanKeyPack(cols.data(), types.data(), (int)keys.size(), mins, mins, pack);
		return 0;
	}
	KeyRanges(cols.data(), types.data(), (int)keys.size(), rows, mins, maxs, scratch, stream);
	if (span_out != nullptr && keys.size() == 1) {
		*span_out = maxs[0] - mins[0] + 1;
	}
	return PlanGroupByRanges(keys, rows, mins, maxs, pack);
}

uint64_t PlanGroupByRanges(const std::vector<DevCol> &keys, uint64_t rows, const uint64_t *mins,
                           const uint64_t *maxs, KeyPack &pack) {
	pack = KeyPack {};
	if (keys.empty() || keys.size() > (size_t)kMaxKeyCols) {
		throw std::runtime_error("gpu_groupby: 1 to 4 group key columns are supported");
	}
	std::vector<const void *> cols;
	std::vector<int> types;
	for (auto &k : keys) {
		if (!TypeIsInteger(k.type)) {
			throw std::runtime_error("gpu_groupby: group key is not an integer/boolean column, which this GPU "
			                         "path requires");
		}
		cols.push_back(k.data);
		types.push_back(k.type);
	}
	if (rows == 0) {
		PlanKeyPack(cols.data(), types.data(), (int)keys.size(), mins, mins, pack);
		return 0;
	}
	if (!PlanKeyPack(cols.data(), types.data(), (int)keys.size(), mins, maxs, pack)) {
		throw std::runtime_error("gpu_groupby: the group keys' value ranges need more than 64 bits once packed");
	}
	// Bound = min(rows, prod(range + 1)), saturating.
	long double product = 1.0L;
	for (size_t c = 0; c < keys.size(); c++) {
		product *= (long double)(maxs[c] - mins[c]) + 1.0L;
	}
	if (product < (long double)rows) {
		return (uint64_t)product;
	}
	return rows;
}

uint64_t GroupByDevice(const std::vector<DevCol> &keys, const std::vector<DevCol> &values,
                       const std::vector<AggReq> &aggs, uint64_t rows, const KeyPack &pack, uint64_t capacity,
                       const OutAlloc &out_alloc, std::vector<void *> &key_out, std::vector<void *> &agg_out,
                       const SortTuning &tuning, Scratch &scratch, hipStream_t stream, bool read_groups,
                       uint64_t *groups_device, const uint64_t *rows_dev, bool narrow_payload) {
	if (aggs.size() > 16) {
		throw std::runtime_error("gpu_groupby: at most 16 aggregates per GROUP BY");
	}
	for (auto &a : aggs) {
		ValidateAggregateReq(a, values);
	}
	for (auto &k : keys) {
		if (k.rows != rows) {
			throw std::runtime_error("gpu_groupby: group key row count mismatch");
		}
	}
	for (auto &v : values) {
		if (v.rows != rows) {
			throw std::runtime_error("gpu_groupby: aggregate input column row count mismatch");
		}
	}

	// Payload = the distinct value columns the aggregates read; components deduplicated per column.
	ReduceSpec spec;
	std::vector<int> payload_of_value(values.size(), -1);
	std::vector<const void *> payload_in;
	std::vector<int> payload_width;
	std::vector<int> payload_type;
	auto payload_for = [&](int value_index) {
		if (payload_of_value[value_index] < 0) {
			if ((int)payload_in.size() >= kMaxPayload) {
				throw std::runtime_error("gpu_groupby: aggregates read more than 8 distinct columns");
			}
			payload_of_value[value_index] = (int)payload_in.size();
			payload_in.push_back(values[value_index].data);
			payload_width.push_back(TypeWidth(values[value_index].type));
			payload_type.push_back(values[value_index].type);
		}
		return payload_of_value[value_index];
	};
	spec.ncomp = 1;
	spec.comp[0].kind = COMP_COUNT;
	auto comp_for = [&](int kind, int value_index) {
		int p = payload_for(value_index);
		for (int c = 0; c < spec.ncomp; c++) {
			if (spec.comp[c].kind == kind && spec.comp[c].payload == p) {
				return c;
			}
		}
		if (spec.ncomp >= kMaxComponents) {
			throw std::runtime_error("gpu_groupby: too many distinct aggregate states (max 8)");
		}
		int c = spec.ncomp++;
		spec.comp[c].kind = kind;
		spec.comp[c].payload = p;
		spec.comp[c].in_type = values[value_index].type;
		return c;
	};
	spec.naggs = (int)aggs.size();
	for (size_t i = 0; i < aggs.size(); i++) {
		const auto &a = aggs[i];
		int comp = 0;
		switch (a.kind) {
		case AGG_COUNT_STAR:
		case AGG_COUNT:
			comp = 0;
			break;
		case AGG_SUM:
		case AGG_AVG:
			comp = comp_for(IsFloat(values[a.input].type) ? COMP_SUM_DD : COMP_SUM_I128, a.input);
			break;
		case AGG_MIN:
			comp = comp_for(IsFloat(values[a.input].type) ? COMP_MIN_F64 : COMP_MIN_I64, a.input);
			break;
		case AGG_MAX:
			comp = comp_for(IsFloat(values[a.input].type) ? COMP_MAX_F64 : COMP_MAX_I64, a.input);
			break;
		}
		spec.aggs[i].kind = a.kind;
		spec.aggs[i].comp = comp;
		spec.aggs[i].out_type = a.out_type;
	}

	// Outputs.
	const uint64_t cap = std::max<uint64_t>(capacity, 1);
	key_out.clear();
	agg_out.clear();
	for (size_t c = 0; c < keys.size(); c++) {
		key_out.push_back(out_alloc(cap * TypeWidth(keys[c].type)));
		spec.key_out[c] = key_out.back();
	}
	for (size_t i = 0; i < aggs.size(); i++) {
		agg_out.push_back(out_alloc(cap * TypeWidth(aggs[i].out_type)));
		spec.aggs[i].out = agg_out.back();
	}
	spec.pack = pack;
	spec.capacity = cap;

	// Word 0: the group count (unless the caller keeps it), word 1: the narrowing overflow flag, so one
	// read-back brings both.
	auto *words = static_cast<uint64_t *>(scratch.Alloc(2 * sizeof(uint64_t)));
	auto *groups_dev = groups_device != nullptr ? groups_device : words;
	auto *overflow = reinterpret_cast<uint32_t *>(words + 1);
	uint32_t narrow = 0;
	if (narrow_payload && read_groups && rows_dev == nullptr && keys.size() == 1 && rows > 0) {
		for (size_t p = 0; p < payload_in.size(); p++) {
			if (payload_type[p] == VT_INT64 && !KnownWide(payload_in[p], rows)) {
				narrow |= 1u << p;
			}
		}
	}
	auto sort_reduce = [&](uint32_t narrow_mask) {
		if (keys.empty()) {
			SortedRun run;
			run.keys = nullptr;
			run.npayload = (int)payload_in.size();
			for (int p = 0; p < run.npayload; p++) {
				run.payload[p] = const_cast<void *>(payload_in[p]);
				run.payload_width[p] = payload_width[p];
			}
			SegmentedReduce(run, rows, spec, groups_dev, scratch, stream, rows_dev);
			return;
		}
		ReduceSpec rs = spec;
		std::vector<int> widths = payload_width;
		SortTuning t = tuning;
		if (narrow_mask != 0) {
			for (int p = 0; p < (int)widths.size(); p++) {
				if (((narrow_mask >> p) & 1u) != 0) {
					widths[p] = 4;
					for (int c = 0; c < rs.ncomp; c++) {
						if (rs.comp[c].kind != COMP_COUNT && rs.comp[c].payload == p) {
							rs.comp[c].in_type = VT_INT32;
						}
					}
				}
			}
			t.narrow = narrow_mask;
			t.overflow = overflow;
			t.zero_also.Add(overflow, 1); // cleared with the sort's look-back state
		}
		t.leave_last = rows_dev == nullptr && RunReduceWorthIt(pack, rs, widths, rows);
		uint32_t *run_state = nullptr;
		if (t.leave_last) {
			t.radix_bits = 8; // the fused pass is an 8-bit digit pass
			const uint64_t words = RunReduceStateWords(rows);
			run_state = static_cast<uint32_t *>(scratch.Alloc(words * sizeof(uint32_t)));
			t.zero_also.Add(run_state, words);
		}
		SortedRun run;
		SortPacked(pack, payload_in.data(), widths.data(), (int)payload_in.size(), rows, t, run, scratch, stream,
		           rows_dev);
		if (run.last_pending) {
			RunReduceLastPass(run, rows, rs, groups_dev, scratch, stream, run_state);
		} else {
			SegmentedReduce(run, rows, rs, groups_dev, scratch, stream, rows_dev);
		}
	};
	if (rows == 0) {
		FillU64(groups_dev, 0ull, 1, stream);
	} else {
		sort_reduce(narrow);
	}
	if (!read_groups) {
		return 0;
	}
	auto read_back = [&](uint64_t &groups, uint64_t &flag) {
		flag = 0;
		if (groups_dev == words) {
			uint64_t host[2] = {};
			ReadBack(host, words, narrow != 0 ? sizeof(host) : sizeof(uint64_t), stream);
			groups = host[0];
			flag = narrow != 0 ? host[1] : 0;
			return;
		}
		VGPU_HIP(hipMemcpyAsync(&groups, groups_dev, sizeof(groups), hipMemcpyDeviceToHost, stream));
		if (narrow != 0) {
			VGPU_HIP(hipMemcpyAsync(&flag, words + 1, sizeof(flag), hipMemcpyDeviceToHost, stream));
		}
		VGPU_HIP(hipStreamSynchronize(stream));
	};
	uint64_t groups = 0, flag = 0;
	read_back(groups, flag);
	if ((uint32_t)flag != 0) {
		// A narrowed column held a value outside INT32: run again carrying every column whole.
		for (size_t p = 0; p < payload_in.size(); p++) {
			if (((narrow >> p) & 1u) != 0) {
				NoteWide(payload_in[p], rows);
			}
		}
		narrow = 0;
		sort_reduce(0);
		read_back(groups, flag);
	}
	return groups;
}

} // namespace vgpu
