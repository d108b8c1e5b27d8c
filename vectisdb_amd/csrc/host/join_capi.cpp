// join_capi.cpp -- extern "C" entry points of the sort-merge join.
#include "../hip/join_kernels.h"
#include "capi_util.h"
#include "device_pool.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

using namespace vgpu;

namespace {

double DdToDouble(uint64_t hi_bits, uint64_t lo_bits) {
	double hi, lo;
	std::memcpy(&hi, &hi_bits, 8);
	std::memcpy(&lo, &lo_bits, 8);
	if (!std::isfinite(hi)) {
		return hi;
	}
	return hi + lo;
}

double I128ToDoubleHost(uint64_t lo, uint64_t hi) {
	__int128 v = (__int128)(((unsigned __int128)hi << 64) | lo);
	return (double)v;
}

} // namespace

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
		for (int32_t q = 0; q < naggs; q++) {
			const auto &v = vals[q];
			const int ot = aggs[q].output_type;
			uint8_t *dst = out + 16 * q;
			double d = 0;
			switch (aggs[q].kind) {
			case AGG_COUNT_STAR:
			case AGG_COUNT:
				std::memcpy(dst, &v.lo, 8);
				break;
			case AGG_SUM:
				if (v.kind == JA_SUM_F64) {
					d = DdToDouble(v.lo, v.hi);
					std::memcpy(dst, &d, 8);
				} else if (ot == VT_HUGEINT) {
					std::memcpy(dst, &v.lo, 8);
					std::memcpy(dst + 8, &v.hi, 8);
				} else if (ot == VT_FLOAT64) {
					d = I128ToDoubleHost(v.lo, v.hi);
					std::memcpy(dst, &d, 8);
				} else {
					std::memcpy(dst, &v.lo, 8);
				}
				break;
			case AGG_AVG: {
				double sum = v.kind == JA_SUM_F64 ? DdToDouble(v.lo, v.hi) : I128ToDoubleHost(v.lo, v.hi);
				double cnt = (double)vals[count_slot].lo;
				d = sum / cnt;
				std::memcpy(dst, &d, 8);
				break;
			}
			default: // MIN / MAX
				std::memcpy(dst, &v.lo, 8);
				break;
			}
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
			VGPU_HIP(hipMemcpyAsync(a_rows, l, total * 4, hipMemcpyDeviceToDevice, s));
			VGPU_HIP(hipMemcpyAsync(b_rows, r, total * 4, hipMemcpyDeviceToDevice, s));
		}
		VGPU_HIP(hipStreamSynchronize(s));
	});
}

} // extern "C"
