// join_op.cpp -- host side of the sort-merge INNER equi-join.
//
// Semantics follow cuda_engine/src/operators/gpu_hash_join.cu:138-267: INNER, exactly one equality key,
// both key columns the same integer/boolean type, output columns selected by name, at most 2^32 output
// rows. The algorithm is sort-merge end to end (both sides radix-sorted, then merged), see join.hip.
#include "../hip/join_kernels.h"
#include "operators.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace vgpu {

namespace {

void ValidateJoinKeys(const DevCol &a, const DevCol &b) {
	if (!TypeIsInteger(a.type) || !TypeIsInteger(b.type)) {
		throw std::runtime_error("gpu_hash_join: join key is not an integer/boolean column, which this GPU path "
		                         "requires");
	}
	if (a.type != b.type) {
		throw std::runtime_error("gpu_hash_join: join key columns must have the same type on both sides");
	}
}

//! One side reduced to its distinct keys: keys (original type, ascending), per-key row counts, optional
//! per-key aggregate columns, and (materialized form) the stable row permutation and run starts.
struct SideRuns {
	uint64_t groups = 0;
	void *keys = nullptr;
	long long *counts = nullptr;
	std::vector<void *> aggs;
	uint32_t *perm = nullptr;
	uint64_t *starts = nullptr;
};

SideRuns ReduceSide(const DevCol &key, const std::vector<DevCol> &values, const std::vector<AggReq> &aggs,
                    bool with_perm, Scratch &scratch, hipStream_t stream) {
	SideRuns s;
	const uint64_t rows = key.rows;
	KeyPack pack;
	const uint32_t *hist8 = nullptr;
	uint64_t bound = PlanGroupBy({key}, rows, pack, scratch, stream, &hist8);
	SortTuning tuning = TuneSort(pack.total_bits, rows);
	tuning.hist8 = hist8;
	if (!with_perm) {
		std::vector<AggReq> all = {{AGG_COUNT_STAR, -1, VT_INT64}};
		all.insert(all.end(), aggs.begin(), aggs.end());
		std::vector<void *> ko, ao;
		OutAlloc alloc = [&](size_t bytes) { return scratch.Alloc(bytes); };
		s.groups = GroupByDevice({key}, values, all, rows, pack, bound, alloc, ko, ao, tuning, scratch, stream, true);
		s.keys = ko[0];
		s.counts = static_cast<long long *>(ao[0]);
		s.aggs.assign(ao.begin() + 1, ao.end());
		return s;
	}
	// Materialized form: sort with a row-id payload, then reduce to (key, count).
	SortedRun run;
	const void *pin[1] = {nullptr};
	int pw[1] = {4};
	SortPacked(pack, pin, pw, 1, rows, tuning, run, scratch, stream);
	s.perm = static_cast<uint32_t *>(run.payload[0]);
	const uint64_t cap = std::max<uint64_t>(bound, 1);
	ReduceSpec spec;
	spec.ncomp = 1;
	spec.comp[0].kind = COMP_COUNT;
	spec.naggs = 1;
	spec.aggs[0].kind = AGG_COUNT_STAR;
	spec.aggs[0].comp = 0;
	spec.aggs[0].out_type = VT_INT64;
	spec.aggs[0].out = scratch.Alloc(cap * 8);
	spec.pack = pack;
	spec.capacity = cap;
	spec.key_out[0] = scratch.Alloc(cap * TypeWidth(key.type));
	// the reduce reads no payload: hide the permutation from it
	SortedRun keys_only = run;
	keys_only.npayload = 0;
	auto *groups_dev = static_cast<uint64_t *>(scratch.Alloc(8));
	SegmentedReduce(keys_only, rows, spec, groups_dev, scratch, stream);
	VGPU_HIP(hipMemcpyAsync(&s.groups, groups_dev, 8, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	s.keys = spec.key_out[0];
	s.counts = static_cast<long long *>(spec.aggs[0].out);
	s.starts = static_cast<uint64_t *>(scratch.Alloc(std::max<uint64_t>(s.groups, 1) * 8));
	ExclusiveScanU64(reinterpret_cast<const uint64_t *>(s.counts), s.starts, s.groups, nullptr, scratch, stream);
	return s;
}

//! First and last key of an ascending distinct-key run, sign-extended.
void KeyBounds(const void *keys, uint64_t groups, int vtype, long long *lo, long long *hi, hipStream_t stream) {
	const int w = TypeWidth(vtype);
	uint8_t buf[2][8] = {};
	VGPU_HIP(hipMemcpyAsync(buf[0], keys, w, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipMemcpyAsync(buf[1], static_cast<const uint8_t *>(keys) + (groups - 1) * w, w, hipMemcpyDeviceToHost,
	                        stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	for (int i = 0; i < 2; i++) {
		long long v = 0;
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
		(i == 0 ? *lo : *hi) = v;
	}
}

//! The double-double state keeps the high double in `lo` and the low-order correction in `hi`.
double DdStateToDouble(const JoinAggValue &v) {
	double head, tail;
	std::memcpy(&head, &v.lo, 8);
	std::memcpy(&tail, &v.hi, 8);
	return std::isfinite(head) ? head + tail : head;
}

double I128StateToDouble(const JoinAggValue &v) {
	__int128 x = (__int128)(((unsigned __int128)v.hi << 64) | v.lo);
	return (double)x;
}

} // namespace

void FinishJoinAggregate(const JoinAggValue &v, int agg_kind, int out_type, uint64_t pairs, void *dst) {
	auto *out = static_cast<uint8_t *>(dst);
	double d = 0;
	switch (agg_kind) {
	case AGG_COUNT_STAR:
	case AGG_COUNT:
		std::memcpy(out, &v.lo, 8);
		return;
	case AGG_SUM:
		if (v.kind == JA_SUM_F64) {
			d = DdStateToDouble(v);
			std::memcpy(out, &d, 8);
		} else if (out_type == VT_HUGEINT) {
			std::memcpy(out, &v.lo, 8);
			std::memcpy(out + 8, &v.hi, 8);
		} else if (out_type == VT_FLOAT64) {
			d = I128StateToDouble(v);
			std::memcpy(out, &d, 8);
		} else {
			std::memcpy(out, &v.lo, 8);
		}
		return;
	case AGG_AVG:
		d = (v.kind == JA_SUM_F64 ? DdStateToDouble(v) : I128StateToDouble(v)) / (double)pairs;
		std::memcpy(out, &d, 8);
		return;
	default: // MIN / MAX keep the 8-byte input value
		std::memcpy(out, &v.lo, 8);
		return;
	}
}

//! Direct-address tables above this size are probed through a key-range partition of the probe side (one
//! XCD's L2 is 4 MB).
constexpr uint64_t kPartitionTableBytes = 4ull << 20;

namespace {

//! A second stream (default priority) on which the probe side is partitioned while the build side is
//! sorted and reduced on the caller's stream, with the two events that order it: `ready` (recorded on the
//! caller's stream, waited by this one) and `done` (the reverse). Streams are pooled per device and reused;
//! the destructor waits for everything issued on the stream, so no exit path can return scratch memory to
//! the pool while the partition still writes it.
class SideStream {
public:
	explicit SideStream(hipStream_t caller) : caller_(caller) {
		VGPU_HIP(hipGetDevice(&device_));
		{
			std::lock_guard<std::mutex> lock(Mutex());
			auto &free_list = Pool()[device_];
			if (!free_list.empty()) {
				stream_ = free_list.back();
				free_list.pop_back();
			}
		}
		if (stream_ == nullptr) {
			VGPU_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
		}
		if (hipEventCreateWithFlags(&ready_, hipEventDisableTiming) != hipSuccess ||
		    hipEventCreateWithFlags(&done_, hipEventDisableTiming) != hipSuccess) {
			Close();
			throw HipError("join: cannot create stream events");
		}
	}
	~SideStream() {
		Close();
	}
	SideStream(const SideStream &) = delete;
	SideStream &operator=(const SideStream &) = delete;

	//! Work issued on stream() after this call starts after everything issued on the caller's stream so far.
	hipStream_t Begin() {
		VGPU_HIP(hipEventRecord(ready_, caller_));
		VGPU_HIP(hipStreamWaitEvent(stream_, ready_, 0));
		return stream_;
	}
	//! Work issued on the caller's stream after this call starts after everything issued on stream().
	void Join() {
		VGPU_HIP(hipEventRecord(done_, stream_));
		VGPU_HIP(hipStreamWaitEvent(caller_, done_, 0));
	}

private:
	static std::mutex &Mutex() {
		static std::mutex m;
		return m;
	}
	static std::map<int, std::vector<hipStream_t>> &Pool() {
		static std::map<int, std::vector<hipStream_t>> pool;
		return pool;
	}
	void Close() {
		if (stream_ != nullptr) {
			const bool ok = hipStreamSynchronize(stream_) == hipSuccess;
			if (ok) {
				std::lock_guard<std::mutex> lock(Mutex());
				Pool()[device_].push_back(stream_);
			}
			stream_ = nullptr; // a stream left in error is not reused
		}
		if (ready_ != nullptr) {
			(void)hipEventDestroy(ready_);
			ready_ = nullptr;
		}
		if (done_ != nullptr) {
			(void)hipEventDestroy(done_);
			done_ = nullptr;
		}
	}

	hipStream_t caller_;
	int device_ = 0;
	hipStream_t stream_ = nullptr;
	hipEvent_t ready_ = nullptr;
	hipEvent_t done_ = nullptr;
};

} // namespace

//! Aggregate join by probing: when every aggregate is a COUNT or an integer SUM/AVG of one column of the
//! smaller (build) side or of one column of the larger (probe) side, the build side is reduced to a
//! direct-address table (per key: row count and payload sum) and the probe side is read once, each row
//! adding its key's count and sum, and its own payload times that count; nothing of the probe side is
//! sorted. Returns false, having written nothing, when the build keys are too sparse for the table or a
//! count or sum does not pack.
bool ProbeJoinAggregate(const DevCol &probe, const DevCol *probe_payload, const DevCol &build, const DevCol *payload,
                        const JoinAggArgs &args, const std::vector<JoinAggReq> &aggs, int build_side,
                        std::vector<JoinAggValue> &result, Scratch &scratch, hipStream_t stream) {
	if (build.type == VT_BOOLEAN || build.rows == 0 || probe.rows == 0) {
		return false;
	}
	if (probe_payload != nullptr && probe_payload->rows != probe.rows) {
		throw std::runtime_error("gpu_hash_join: aggregate input row count mismatch");
	}
	// The build keys' range decides whether a direct-address table is possible at all.
	const uint64_t rows = build.rows;
	auto *h8 = static_cast<uint32_t *>(scratch.Alloc((8 * 256 + 4) * sizeof(uint32_t)));
	uint64_t omin = 0, omax = 0;
	KeyRangeHist8(build.data, build.type, rows, &omin, &omax, h8, stream);
	const uint64_t span = omax - omin + 1;
	if (span == 0 || span > 2 * rows + (1ull << 20) || rows >= (1ull << 30)) {
		return false;
	}
	const uint64_t flip = TypeIsSigned(build.type) ? (1ull << 63) : 0ull;
	std::unique_ptr<JoinIndex> index;
	const uint64_t *dense = nullptr;
	uint64_t dense_min = omin ^ flip, dense_span = span;
	if (payload != nullptr && payload->rows != rows) {
		throw std::runtime_error("gpu_hash_join: aggregate input row count mismatch");
	}
	// The table is larger than one XCD's L2 and every line of it is read several times (or the probe payload
	// has to travel with its key): the probe keys are grouped by key range first. The grouping needs only the
	// key range, so it runs on a second stream while the build side is sorted and reduced on this one.
	const bool partition =
	    span <= (1ull << 32) && ((span * 8 > kPartitionTableBytes && probe.rows >= span) || probe_payload != nullptr);
	if (!partition && probe_payload != nullptr) {
		return false;
	}
	std::unique_ptr<SideStream> side;
	ProbePartition part;
	if (partition) {
		part = PlanProbePartition(probe.type, probe.rows, probe_payload != nullptr,
		                          probe_payload ? TypeWidth(probe_payload->type) : 0, dense_min, span, scratch);
		side.reset(new SideStream(stream));
		LaunchProbePartition(part, probe.data, probe_payload ? probe_payload->data : nullptr, side->Begin());
	}
	// The table: the build side sorted (one key column, payload alongside) and each key's run turned into its
	// entry in one pass (no per-key arrays); not built when the build keys span so many values that even
	// all-distinct keys could not fill half of it. A key with more than 64 build rows sends the build through
	// the GROUP BY operator instead. With the partition, the table's 16-bit form is made in the same go, and
	// one read-back returns both flags.
	const uint16_t *dense16 = nullptr;
	{
		KeyPack pack;
		PlanKeyPackAligned(build.data, build.type, omin, omax, pack);
		SortTuning tuning = TuneSort(pack.total_bits, rows);
		if (tuning.radix_bits == 8) {
			tuning.hist8 = h8; // byte digits of the aligned packing are the key's bytes
		}
		const void *pin[1] = {payload != nullptr ? payload->data : nullptr};
		int pw[1] = {payload != nullptr ? TypeWidth(payload->type) : 0};
		SortedRun run;
		SortPacked(pack, pin, pw, payload != nullptr ? 1 : 0, rows, tuning, run, scratch, stream);
		auto *table = static_cast<uint64_t *>(scratch.Alloc(span * 8));
		auto *flags = static_cast<unsigned int *>(scratch.Alloc(2 * 4));
		LaunchDenseFromRun(run.keys, run.keys64, payload != nullptr ? run.payload[0] : nullptr, pw[0], rows,
		                   pack.min[0], flip, dense_min, span, table, flags, stream);
		if (partition) {
			dense16 = LaunchNarrowDense(table, span, flags + 1, scratch, stream);
		}
		unsigned int f[2] = {0u, 1u};
		VGPU_HIP(hipMemcpyAsync(f, flags, (partition ? 2 : 1) * 4, hipMemcpyDeviceToHost, stream));
		VGPU_HIP(hipStreamSynchronize(stream));
		if ((f[0] & 1u) != 0) {
			return false; // an entry does not pack
		}
		dense = table;
		if (f[1] != 0) {
			dense16 = nullptr; // an entry does not fit 16 bits
		}
		if ((f[0] & 2u) != 0) {
			DevCol pay {};
			if (payload != nullptr) {
				pay = *payload;
			}
			index = BuildJoinIndex(build, pay, stream, true);
			if (index->dense == nullptr) {
				return false;
			}
			dense = index->dense;
			if (partition && (index->dense_min != dense_min || index->dense_span != dense_span)) {
				// the index spans other keys than the key-range pass found: group the probe side again
				part = PlanProbePartition(probe.type, probe.rows, probe_payload != nullptr,
				                          probe_payload ? TypeWidth(probe_payload->type) : 0, index->dense_min,
				                          index->dense_span, scratch);
				LaunchProbePartition(part, probe.data, probe_payload ? probe_payload->data : nullptr, stream);
			}
			dense_min = index->dense_min;
			dense_span = index->dense_span;
			dense16 = nullptr;
			if (partition) {
				dense16 = LaunchNarrowDense(dense, dense_span, flags + 1, scratch, stream);
				VGPU_HIP(hipMemcpyAsync(&f[1], flags + 1, 4, hipMemcpyDeviceToHost, stream));
				VGPU_HIP(hipStreamSynchronize(stream));
				if (f[1] != 0) {
					dense16 = nullptr;
				}
			}
		}
	}
	auto *acc = static_cast<uint64_t *>(scratch.Alloc(5 * 8));
	if (partition) {
		side->Join(); // the probe reads the partition
		LaunchPartitionedProbe(part, dense, dense16, acc, stream);
	} else {
		const uint64_t row0 = 0, len = probe.rows;
		const DeviceRanges r = UploadRanges(&row0, &len, 1, probe.rows, scratch, stream);
		RangeProbeAgg(probe.type, probe.data, r.seg, r.rows, r.lens, 1, r.nvirt, nullptr, nullptr, nullptr, 0, dense,
		              dense_min, dense_span, acc, stream);
	}
	uint64_t host[5] = {};
	// RangeProbeAgg fills 3 words; the partitioned probe 5 (the probe-payload sum last)
	VGPU_HIP(hipMemcpyAsync(host, acc, (probe_payload != nullptr ? 5 : 3) * 8, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	for (int q = 0; q < args.naggs; q++) {
		result[q].kind = args.kind[q];
		if (args.kind[q] == JA_COUNT) {
			result[q].lo = host[0];
			result[q].hi = 0;
		} else if (aggs[q].side == build_side) {
			result[q].lo = host[1];
			result[q].hi = host[2];
		} else {
			result[q].lo = host[3];
			result[q].hi = host[4];
		}
	}
	return true;
}

std::vector<JoinAggValue> JoinAggregateDevice(const DevCol &akey, const std::vector<DevCol> &avals, const DevCol &bkey,
                                              const std::vector<DevCol> &bvals, const std::vector<JoinAggReq> &aggs,
                                              Scratch &scratch, hipStream_t stream) {
	ValidateJoinKeys(akey, bkey);
	if (aggs.size() > (size_t)kJoinAggMax) {
		throw std::runtime_error("gpu_hash_join: at most 8 aggregates over a join");
	}
	// Per-side aggregate requests.
	std::vector<AggReq> side_aggs[2];
	std::vector<int> side_slot(aggs.size(), -1);
	JoinAggArgs args;
	args.naggs = (int)aggs.size();
	for (size_t q = 0; q < aggs.size(); q++) {
		const auto &a = aggs[q];
		const std::vector<DevCol> &vals = a.side == 0 ? avals : bvals;
		if (a.kind == AGG_COUNT_STAR || a.kind == AGG_COUNT) {
			args.kind[q] = JA_COUNT;
			continue;
		}
		if (a.side != 0 && a.side != 1) {
			throw std::runtime_error("gpu_hash_join: aggregate side must be 0 or 1");
		}
		if (a.input < 0 || a.input >= (int)vals.size()) {
			throw std::runtime_error("gpu_hash_join: aggregate input out of range");
		}
		const int in = vals[a.input].type;
		const bool flt = in == VT_FLOAT32 || in == VT_FLOAT64;
		AggReq r;
		r.input = a.input;
		args.side[q] = a.side;
		switch (a.kind) {
		case AGG_SUM:
		case AGG_AVG:
			r.kind = AGG_SUM;
			r.out_type = flt ? VT_FLOAT64 : VT_HUGEINT;
			args.kind[q] = flt ? JA_SUM_F64 : JA_SUM_I128;
			break;
		case AGG_MIN:
		case AGG_MAX:
			if (!(in == VT_INT64 || in == VT_FLOAT64)) {
				throw std::runtime_error("gpu_hash_join: MIN/MAX over a join supports BIGINT and DOUBLE inputs");
			}
			r.kind = a.kind;
			r.out_type = in;
			args.kind[q] = flt ? (a.kind == AGG_MIN ? JA_MIN_F64 : JA_MAX_F64)
			                   : (a.kind == AGG_MIN ? JA_MIN_I64 : JA_MAX_I64);
			break;
		default:
			throw std::runtime_error("gpu_hash_join: unsupported aggregate over a join");
		}
		side_slot[q] = (int)side_aggs[a.side].size();
		side_aggs[a.side].push_back(r);
	}
	std::vector<JoinAggValue> result(aggs.size());
	if (akey.rows == 0 || bkey.rows == 0) {
		return result; // zero rows: COUNT 0, SUM 0 (the caller applies SQL NULL rules for empty input)
	}
	// A side that only contributes its row counts and is the larger one is first narrowed to the other
	// side's key range: its rows outside that range have no partner, and the narrower key span can save a
	// radix pass (a probe side over twice the build keys' range loses a whole digit).
	int narrow = -1;
	if (akey.type != VT_BOOLEAN) {
		if (side_aggs[0].empty() && akey.rows >= bkey.rows) {
			narrow = 0;
		} else if (side_aggs[1].empty() && bkey.rows > akey.rows) {
			narrow = 1;
		}
	}
	// VGPU_JOIN_FORM=sort keeps every aggregate on the sort-merge form (to measure it beside the probe form).
	const char *form = std::getenv("VGPU_JOIN_FORM");
	const bool sort_only = form != nullptr && std::strcmp(form, "sort") == 0;
	if (akey.type != VT_BOOLEAN && !sort_only) {
		// Probe form: the larger side is read once against a table of the smaller one. Every aggregate must
		// be a COUNT or an integer SUM/AVG of one column of either side.
		const int probe_side = akey.rows >= bkey.rows ? 0 : 1;
		const int build_side = 1 - probe_side;
		int input[2] = {-1, -1};
		bool probe_ok = true;
		for (size_t q = 0; q < aggs.size() && probe_ok; q++) {
			if (args.kind[q] == JA_COUNT) {
				continue;
			}
			const int sd = aggs[q].side;
			probe_ok = args.kind[q] == JA_SUM_I128 && (sd == 0 || sd == 1) &&
			           (input[sd] < 0 || input[sd] == aggs[q].input);
			if (probe_ok) {
				input[sd] = aggs[q].input;
			}
		}
		if (probe_ok) {
			const std::vector<DevCol> *vals[2] = {&avals, &bvals};
			const DevCol *build_payload = input[build_side] >= 0 ? &(*vals[build_side])[input[build_side]] : nullptr;
			const DevCol *probe_payload = input[probe_side] >= 0 ? &(*vals[probe_side])[input[probe_side]] : nullptr;
			if (ProbeJoinAggregate(probe_side == 0 ? akey : bkey, probe_payload, build_side == 0 ? akey : bkey,
			                       build_payload, args, aggs, build_side, result, scratch, stream)) {
				return result;
			}
		}
	}
	SideRuns A, B;
	if (narrow < 0) {
		A = ReduceSide(akey, avals, side_aggs[0], false, scratch, stream);
		B = ReduceSide(bkey, bvals, side_aggs[1], false, scratch, stream);
	} else {
		const DevCol &wide = narrow == 0 ? akey : bkey;
		SideRuns &other = narrow == 0 ? B : A;
		other = narrow == 0 ? ReduceSide(bkey, bvals, side_aggs[1], false, scratch, stream)
		                    : ReduceSide(akey, avals, side_aggs[0], false, scratch, stream);
		if (other.groups == 0) {
			return result;
		}
		long long lo = 0, hi = 0;
		KeyBounds(other.keys, other.groups, wide.type, &lo, &hi, stream);
		DevCol kept = wide;
		void *buf = scratch.Alloc(wide.rows * TypeWidth(wide.type));
		kept.data = buf;
		kept.rows = KeepKeyRange(wide.type, wide.data, wide.rows, lo, hi, buf, scratch, stream);
		if (kept.rows == 0) {
			return result;
		}
		(narrow == 0 ? A : B) = ReduceSide(kept, {}, {}, false, scratch, stream);
	}
	auto *match = static_cast<long long *>(scratch.Alloc(std::max<uint64_t>(A.groups, 1) * 8));
	MergeMatch(A.keys, A.groups, B.keys, B.groups, akey.type, match, stream);
	args.na = A.groups;
	args.match = match;
	args.ca = A.counts;
	args.cb = B.counts;
	for (size_t q = 0; q < aggs.size(); q++) {
		if (side_slot[q] >= 0) {
			args.per_key[q] = (aggs[q].side == 0 ? A : B).aggs[side_slot[q]];
		}
	}
	auto *res_dev = static_cast<uint64_t *>(scratch.Alloc(kJoinAggMax * 2 * 8));
	JoinAggregate(args, res_dev, scratch, stream);
	uint64_t host[kJoinAggMax * 2];
	VGPU_HIP(hipMemcpyAsync(host, res_dev, sizeof(host), hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	for (size_t q = 0; q < aggs.size(); q++) {
		result[q].lo = host[2 * q];
		result[q].hi = host[2 * q + 1];
		result[q].kind = args.kind[q];
	}
	return result;
}

uint64_t JoinMaterializeDevice(const DevCol &akey, const DevCol &bkey, uint32_t **left_rows, uint32_t **right_rows,
                               Scratch &scratch, hipStream_t stream) {
	ValidateJoinKeys(akey, bkey);
	*left_rows = nullptr;
	*right_rows = nullptr;
	if (akey.rows == 0 || bkey.rows == 0) {
		return 0;
	}
	SideRuns A = ReduceSide(akey, {}, {}, true, scratch, stream);
	SideRuns B = ReduceSide(bkey, {}, {}, true, scratch, stream);
	auto *match = static_cast<long long *>(scratch.Alloc(std::max<uint64_t>(A.groups, 1) * 8));
	MergeMatch(A.keys, A.groups, B.keys, B.groups, akey.type, match, stream);
	auto *pairs = static_cast<uint64_t *>(scratch.Alloc(std::max<uint64_t>(A.groups, 1) * 8));
	PairCounts(match, A.counts, B.counts, A.groups, pairs, stream);
	auto *offsets = static_cast<uint64_t *>(scratch.Alloc(std::max<uint64_t>(A.groups, 1) * 8));
	auto *total_dev = static_cast<uint64_t *>(scratch.Alloc(8));
	ExclusiveScanU64(pairs, offsets, A.groups, total_dev, scratch, stream);
	uint64_t total = 0;
	VGPU_HIP(hipMemcpyAsync(&total, total_dev, 8, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	constexpr uint64_t kMaxJoinOutputRows = 1ull << 32;
	if (total > kMaxJoinOutputRows) {
		throw std::runtime_error("gpu_hash_join: join result of " + std::to_string(total) +
		                         " rows is too large for GPU execution");
	}
	if (total == 0) {
		return 0;
	}
	*left_rows = static_cast<uint32_t *>(scratch.Alloc(total * 4));
	*right_rows = static_cast<uint32_t *>(scratch.Alloc(total * 4));
	ExpandPairs(offsets, A.groups, match, B.counts, reinterpret_cast<const long long *>(A.starts),
	            reinterpret_cast<const long long *>(B.starts), A.perm, B.perm, total, *left_rows, *right_rows, stream);
	return total;
}

} // namespace vgpu
