// capi.cpp -- extern "C" entry points declared in include/vgpu.h. Each wraps a C++ operator in a
// try/catch so no exception crosses the C boundary; errors land in a thread-local message.
#include "../../../include/vgpu.h"
#include "capi_util.h"
#include "device_pool.h"
#include "operators.h"

#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace vgpu {
namespace {
thread_local std::string g_last_error;
} // namespace

void SetLastError(const std::string &message) {
	g_last_error = message;
}

std::vector<DevCol> ToDevCols(const vgpu_column *cols, int32_t n) {
	std::vector<DevCol> out;
	for (int32_t i = 0; i < n; i++) {
		DevCol c;
		c.data = cols[i].data;
		c.type = cols[i].type;
		c.rows = cols[i].rows;
		out.push_back(c);
	}
	return out;
}

BatchedGroupByArgs MakeBatchedArgs(const vgpu_column *cols, int32_t ncols, uint64_t rows, const uint64_t *offsets,
                                   int32_t nsegments, const vgpu_expr *preds, int32_t npreds, int32_t key_col,
                                   const int32_t *value_cols, int32_t nvalues, const vgpu_agg *aggs, int32_t naggs,
                                   void *segment_out, void *key_out, void *const *agg_out, uint64_t capacity,
                                   uint64_t *segment_group_offsets) {
	if (ncols < 0 || nsegments <= 0 || nvalues < 0 || naggs < 0 || npreds < 0) {
		throw std::runtime_error("batched group by: bad counts");
	}
	BatchedGroupByArgs a;
	a.cols = ToDevCols(cols, ncols);
	a.rows = rows;
	a.seg_offsets_dev = offsets;
	a.nseg = static_cast<uint32_t>(nsegments);
	for (int32_t i = 0; i < npreds; i++) {
		if (preds[i].body == nullptr) {
			throw std::runtime_error("batched group by: predicate body is NULL");
		}
		FusedExpr f;
		f.body = preds[i].body;
		f.out_type = preds[i].output_type;
		f.slots.assign(preds[i].inputs, preds[i].inputs + preds[i].ninputs);
		a.preds.push_back(f);
	}
	a.key_col = key_col;
	a.value_cols.assign(value_cols, value_cols + nvalues);
	for (int32_t i = 0; i < naggs; i++) {
		if (aggs[i].kind != AGG_COUNT_STAR && (aggs[i].input < 0 || aggs[i].input >= nvalues)) {
			throw std::runtime_error("batched group by: aggregate input out of range");
		}
		a.aggs.push_back({aggs[i].kind, aggs[i].input, aggs[i].output_type});
	}
	a.seg_out = segment_out;
	a.key_out = key_out;
	a.agg_out.assign(agg_out, agg_out + naggs);
	a.capacity = capacity;
	a.seg_group_offsets_host = segment_group_offsets;
	return a;
}

} // namespace vgpu

using namespace vgpu;

extern "C" {

const char *vgpu_last_error(void) {
	return g_last_error.c_str();
}

int vgpu_abi_version(void) {
	return 1;
}

int vgpu_device_init(int device) {
	return Guard([&] {
		VGPU_HIP(hipSetDevice(device));
		VGPU_HIP(hipFree(nullptr));
		NoAllocScratchInit();
	});
}

int vgpu_group_by_bound(const vgpu_column *keys, int32_t nkeys, uint64_t rows, uint64_t *bound, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		KeyPack pack;
		*bound = PlanGroupBy(ToDevCols(keys, nkeys), rows, pack, arena, s);
	});
}

int vgpu_group_by(const vgpu_column *keys, int32_t nkeys, const vgpu_column *values, int32_t nvalues,
                  const vgpu_agg *aggs, int32_t naggs, uint64_t rows, void *const *key_out, void *const *agg_out,
                  uint64_t capacity, uint64_t *groups, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		auto kc = ToDevCols(keys, nkeys);
		auto vc = ToDevCols(values, nvalues);
		std::vector<AggReq> reqs;
		for (int32_t i = 0; i < naggs; i++) {
			reqs.push_back({aggs[i].kind, aggs[i].input, aggs[i].output_type});
		}
		KeyPack pack;
		const uint32_t *hist8 = nullptr;
		uint64_t bound = PlanGroupBy(kc, rows, pack, arena, s, &hist8);
		if (bound > capacity) {
			throw std::runtime_error("vgpu_group_by: output capacity " + std::to_string(capacity) +
			                         " is below the group bound " + std::to_string(bound));
		}
		// Outputs are the caller's buffers, handed out in order.
		std::vector<void *> fixed;
		for (int32_t c = 0; c < nkeys; c++) {
			fixed.push_back(key_out[c]);
		}
		for (int32_t i = 0; i < naggs; i++) {
			fixed.push_back(agg_out[i]);
		}
		size_t next = 0;
		OutAlloc alloc = [&](size_t) { return fixed.at(next++); };
		std::vector<void *> ko, ao;
		SortTuning tuning = TuneSort(pack.total_bits, rows);
		tuning.hist8 = hist8;
		*groups = GroupByDevice(kc, vc, reqs, rows, pack, capacity, alloc, ko, ao, tuning, arena, s, true, nullptr,
		                        nullptr, true);
	});
}


int vgpu_batched_filter_group_by(const vgpu_column *cols, int32_t ncols, uint64_t rows,
                                 const uint64_t *segment_offsets, int32_t nsegments, const vgpu_expr *preds,
                                 int32_t npreds, int32_t key_col, const int32_t *value_cols, int32_t nvalues,
                                 const vgpu_agg *aggs, int32_t naggs, void *segment_out, void *key_out,
                                 void *const *agg_out, uint64_t capacity, uint64_t *groups,
                                 uint64_t *segment_group_offsets, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		auto a = MakeBatchedArgs(cols, ncols, rows, segment_offsets, nsegments, preds, npreds, key_col, value_cols,
		                         nvalues, aggs, naggs, segment_out, key_out, agg_out, capacity, segment_group_offsets);
		*groups = BatchedFilterGroupByDevice(a, arena, s);
	});
}

int vgpu_per_query_filter_group_by(const vgpu_column *cols, int32_t ncols, uint64_t rows,
                                   const uint64_t *segment_offsets_host, int32_t nsegments, const vgpu_expr *preds,
                                   int32_t npreds, int32_t key_col, const int32_t *value_cols, int32_t nvalues,
                                   const vgpu_agg *aggs, int32_t naggs, void *segment_out, void *key_out,
                                   void *const *agg_out, uint64_t capacity, uint64_t *groups,
                                   uint64_t *segment_group_offsets, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		auto a = MakeBatchedArgs(cols, ncols, rows, nullptr, nsegments, preds, npreds, key_col, value_cols, nvalues,
		                         aggs, naggs, segment_out, key_out, agg_out, capacity, segment_group_offsets);
		for (auto &c : a.cols) {
			if (c.rows != rows) {
				throw std::runtime_error("per-query group by: column row count mismatch");
			}
		}
		if (segment_offsets_host[nsegments] > rows) {
			throw std::runtime_error("per-query group by: row offsets past the table");
		}
		*groups = PerQueryFilterGroupByDevice(a, segment_offsets_host, arena, s);
	});
}

int vgpu_sort_column(const vgpu_column *key, void *keys_out, uint32_t *perm_out, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		if (!TypeIsInteger(key->type)) {
			throw std::runtime_error("vgpu_sort_column: integer/boolean key required");
		}
		const uint64_t n = key->rows;
		if (n == 0) {
			return;
		}
		const void *cols[1] = {key->data};
		int types[1] = {key->type};
		uint64_t mn[1], mx[1];
		KeyRanges(cols, types, 1, n, mn, mx, arena, s);
		KeyPack pack;
		PlanKeyPack(cols, types, 1, mn, mx, pack);
		const void *pin[2] = {nullptr, key->data};
		int pw[2] = {4, TypeWidth(key->type)};
		SortedRun run;
		SortPacked(pack, pin, pw, 2, n, TuneSort(pack.total_bits, n), run, arena, s);
		CopyDevice(keys_out, run.payload[1], n * pw[1], s);
		if (perm_out != nullptr) {
			CopyDevice(perm_out, run.payload[0], n * 4, s);
		}
	});
}

int vgpu_filter_gt_i32(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_host,
                       void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		auto *count = static_cast<uint64_t *>(arena.Alloc(sizeof(uint64_t)));
		FilterGtI32(in, n, threshold, out, count, arena, s);
		VGPU_HIP(hipMemcpyAsync(count_host, count, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
		VGPU_HIP(hipStreamSynchronize(s));
	});
}

int vgpu_filter_gt_i32_noalloc(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_host,
                               void *stream) {
	return Guard([&] { FilterGtI32NoAlloc(in, n, threshold, out, count_host, static_cast<hipStream_t>(stream)); });
}

uint64_t vgpu_filter_gt_i32_noalloc_max_rows(void) {
	return FilterGtI32NoAllocMaxRows();
}

int vgpu_gather(const void *src, int32_t width, const void *idx, int32_t idx64, uint64_t n, void *out,
                void *stream) {
	return Guard([&] { Gather(src, width, idx, idx64 != 0, n, out, static_cast<hipStream_t>(stream)); });
}

int vgpu_generate_column(void *out, int32_t type, uint64_t n, int32_t dist, int64_t lo, int64_t hi, uint64_t seed,
                         uint64_t stream_id, const double *cdf, uint64_t cdf_len, const int64_t *perm,
                         void *stream) {
	return Guard([&] {
		GenerateColumn(out, type, n, dist, (long long)lo, (long long)hi, seed, stream_id, cdf, cdf_len,
		               reinterpret_cast<const long long *>(perm), static_cast<hipStream_t>(stream));
	});
}

int vgpu_device_copy(void *dst, const void *src, uint64_t bytes, void *stream) {
	return Guard([&] { CopyDevice(dst, src, bytes, static_cast<hipStream_t>(stream)); });
}

int vgpu_profile_enable(int32_t on) {
	return Guard([&] { ProfileEnable(on != 0); });
}

int vgpu_profile_only(const char *name) {
	return Guard([&] { ProfileOnly(name); });
}

int vgpu_profile_reset(void) {
	return Guard([&] { ProfileReset(); });
}

int vgpu_profile_read(char *buf, size_t cap, size_t *needed) {
	return Guard([&] {
		auto entries = ProfileSnapshot();
		std::ostringstream os;
		os << "[";
		for (size_t i = 0; i < entries.size(); i++) {
			char line[512];
			std::snprintf(line, sizeof(line), "%s{\"name\":\"%s\",\"ms\":%.6f,\"launches\":%llu,\"bytes\":%.1f}",
			              i ? "," : "", entries[i].name.c_str(), entries[i].total_ms,
			              (unsigned long long)entries[i].launches, entries[i].bytes);
			os << line;
		}
		os << "]";
		std::string text = os.str();
		if (needed != nullptr) {
			*needed = text.size() + 1;
		}
		if (buf == nullptr || cap < text.size() + 1) {
			throw std::runtime_error("vgpu_profile_read: buffer too small");
		}
		std::memcpy(buf, text.c_str(), text.size() + 1);
	});
}

int vgpu_sort_autotune(int32_t on) {
	return Guard([&] { SortAutotuneEnable(on != 0); });
}

int vgpu_group_by_fused_last_pass(int32_t mode) {
	return Guard([&] {
		if (mode < 0 || mode > 2) {
			throw std::runtime_error("vgpu_group_by_fused_last_pass: mode is 0, 1 or 2");
		}
		SetRunReduceMode(mode);
	});
}

int vgpu_sort_autotune_report(char *buf, size_t cap, size_t *needed) {
	return Guard([&] {
		const std::string r = SortAutotuneReport();
		if (needed != nullptr) {
			*needed = r.size() + 1;
		}
		if (buf == nullptr || cap < r.size() + 1) {
			throw std::runtime_error("sort autotune report: buffer too small");
		}
		std::memcpy(buf, r.c_str(), r.size() + 1);
	});
}

int vgpu_pool_stats(uint64_t *in_use, uint64_t *cached, uint64_t *peak) {
	return Guard([&] {
		auto &pool = DevicePool::Instance();
		*in_use = pool.BytesInUse();
		*cached = pool.BytesCached();
		*peak = pool.PeakBytesInUse();
	});
}

int vgpu_pool_reset_peak(void) {
	return Guard([&] { DevicePool::Instance().ResetPeak(); });
}

int vgpu_pool_trim(void) {
	return Guard([&] { DevicePool::Instance().Trim(); });
}

} // extern "C"
