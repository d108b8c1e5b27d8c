// join_capi.cpp -- extern "C" entry points of the sort-merge join.
#include "../hip/join_kernels.h"
#include "capi_util.h"
#include "device_pool.h"

#include <cstring>
#include <stdexcept>
#include <string>

using namespace vgpu;

extern "C" {

int vgpu_join_aggregate(const vgpu_column *akey, const vgpu_column *avals, int32_t navals, const vgpu_column *bkey,
                        const vgpu_column *bvals, int32_t nbvals, const vgpu_join_agg *aggs, int32_t naggs,
                        void *results, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		std::vector<JoinAggReq> reqs;
		bool need_count = false;
		for (int32_t q = 0; q < naggs; q++) {
			JoinAggReq r;
			r.kind = aggs[q].kind;
			r.side = aggs[q].side;
			r.input = aggs[q].input;
			r.out_type = aggs[q].output_type;
			need_count = need_count || r.kind == AGG_AVG;
			reqs.push_back(r);
		}
		size_t count_slot = reqs.size();
		if (need_count) {
			reqs.push_back(JoinAggReq {AGG_COUNT_STAR, -1, -1, VT_INT64});
		}
		auto vals = JoinAggregateDevice(ToDevCols(akey, 1)[0], ToDevCols(avals, navals), ToDevCols(bkey, 1)[0],
		                                ToDevCols(bvals, nbvals), reqs, arena, s);
		auto *out = static_cast<uint8_t *>(results);
		std::memset(out, 0, (size_t)naggs * 16);
		const uint64_t pairs = need_count ? vals[count_slot].lo : 0;
		for (int32_t q = 0; q < naggs; q++) {
			FinishJoinAggregate(vals[q], aggs[q].kind, aggs[q].output_type, pairs, out + 16 * q);
		}
	});
}

int vgpu_join_pairs(const vgpu_column *akey, const vgpu_column *bkey, uint32_t *a_rows, uint32_t *b_rows,
                    uint64_t capacity, uint64_t *count, void *stream) {
	return Guard([&] {
		auto s = static_cast<hipStream_t>(stream);
		PoolArena arena(s);
		uint32_t *l = nullptr, *r = nullptr;
		uint64_t total = JoinMaterializeDevice(ToDevCols(akey, 1)[0], ToDevCols(bkey, 1)[0], &l, &r, arena, s);
		*count = total;
		if (total > capacity) {
			throw std::runtime_error("vgpu_join_pairs: " + std::to_string(total) + " pairs exceed capacity " +
			                         std::to_string(capacity));
		}
		if (total > 0) {
			CopyDevice(a_rows, l, total * 4, s);
			CopyDevice(b_rows, r, total * 4, s);
		}
		VGPU_HIP(hipStreamSynchronize(s));
	});
}

} // extern "C"
