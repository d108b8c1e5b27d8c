// join_kernels.h -- aggregate combine of the sort-merge join: argument block, per-aggregate state
// (two 64-bit words), and the launch interface of join.hip.
#pragma once

#include "kernels.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgpu {

constexpr int kJoinAggMax = 8;

enum JoinAggKind : int {
	JA_COUNT = 0,   // sum over matched keys of ca * cb            (uint128)
	JA_SUM_I128 = 1, // side's per-key int128 sum x other side's count (int128)
	JA_SUM_F64 = 2,  // side's per-key double sum x other side's count (double-double)
	JA_MIN_I64 = 3,
	JA_MAX_I64 = 4,
	JA_MIN_F64 = 5,
	JA_MAX_F64 = 6,
};

struct JoinAggArgs {
	uint64_t na = 0;
	const long long *match = nullptr; // per A key: index of the equal B key, or -1
	const long long *ca = nullptr;    // rows per A key
	const long long *cb = nullptr;    // rows per B key
	int naggs = 0;
	int kind[kJoinAggMax] = {};
	int side[kJoinAggMax] = {};           // 0 = A, 1 = B
	const void *per_key[kJoinAggMax] = {}; // that side's per-key partial (HUGEINT / double / int64)
};

#ifdef __HIPCC__
__device__ __forceinline__ void JoinAggIdentity(const JoinAggArgs &a, int q, uint64_t &x, uint64_t &y) {
	x = 0;
	y = 0;
	if (q >= a.naggs) {
		return;
	}
	switch (a.kind[q]) {
	case JA_MIN_I64:
		x = (uint64_t)0x7fffffffffffffffll;
		break;
	case JA_MAX_I64:
		x = (uint64_t)0x8000000000000000ull;
		break;
	case JA_MIN_F64:
		x = (uint64_t)__double_as_longlong(__builtin_nan(""));
		break;
	case JA_MAX_F64:
		x = (uint64_t)__double_as_longlong(-__builtin_inf());
		break;
	default:
		break;
	}
}

__device__ __forceinline__ bool JaLess(double x, double y) {
	if (isnan(y)) {
		return !isnan(x);
	}
	return !isnan(x) && x < y;
}

__device__ __forceinline__ void JaAdd128(uint64_t &lo, uint64_t &hi, uint64_t lo2, uint64_t hi2) {
	uint64_t n = lo + lo2;
	hi = hi + hi2 + (n < lo ? 1ull : 0ull);
	lo = n;
}

__device__ __forceinline__ void JaDdAdd(uint64_t &xs, uint64_t &ys, double hi2, double lo2) {
	double hi = __longlong_as_double((long long)xs), lo = __longlong_as_double((long long)ys);
	double s = hi + hi2;
	double bb = s - hi;
	double e = (hi - (s - bb)) + (hi2 - bb);
	e += lo + lo2;
	hi = s + e;
	lo = e - (hi - s);
	xs = (uint64_t)__double_as_longlong(hi);
	ys = (uint64_t)__double_as_longlong(lo);
}

__device__ __forceinline__ void JoinAggAccumulate(const JoinAggArgs &a, uint64_t i, uint64_t m, uint64_t *acc) {
	const uint64_t ca = (uint64_t)a.ca[i], cb = (uint64_t)a.cb[m];
#pragma unroll
	for (int q = 0; q < kJoinAggMax; q++) {
		if (q >= a.naggs) {
			continue;
		}
		const uint64_t own = a.side[q] == 0 ? i : m;
		const uint64_t other = a.side[q] == 0 ? cb : ca;
		uint64_t &x = acc[2 * q];
		uint64_t &y = acc[2 * q + 1];
		switch (a.kind[q]) {
		case JA_COUNT: {
			uint64_t lo = ca * cb, hi = __umul64hi(ca, cb);
			JaAdd128(x, y, lo, hi);
			break;
		}
		case JA_SUM_I128: {
			const uint64_t *s = (const uint64_t *)a.per_key[q];
			uint64_t slo = s[2 * own], shi = s[2 * own + 1];
			uint64_t plo = slo * other;
			uint64_t phi = shi * other + __umul64hi(slo, other); // two's complement product mod 2^128
			JaAdd128(x, y, plo, phi);
			break;
		}
		case JA_SUM_F64: {
			double v = ((const double *)a.per_key[q])[own];
			double c = (double)other;
			double p = v * c;
			double e = __fma_rn(v, c, -p);
			JaDdAdd(x, y, p, e);
			break;
		}
		case JA_MIN_I64: {
			long long v = ((const long long *)a.per_key[q])[own];
			if (v < (long long)x) {
				x = (uint64_t)v;
			}
			break;
		}
		case JA_MAX_I64: {
			long long v = ((const long long *)a.per_key[q])[own];
			if (v > (long long)x) {
				x = (uint64_t)v;
			}
			break;
		}
		case JA_MIN_F64: {
			double v = ((const double *)a.per_key[q])[own];
			if (JaLess(v, __longlong_as_double((long long)x))) {
				x = (uint64_t)__double_as_longlong(v);
			}
			break;
		}
		case JA_MAX_F64: {
			double v = ((const double *)a.per_key[q])[own];
			if (JaLess(__longlong_as_double((long long)x), v)) {
				x = (uint64_t)__double_as_longlong(v);
			}
			break;
		}
		default:
			break;
		}
	}
}

//! acc = acc (earlier) merged with t (later).
__device__ __forceinline__ void JoinAggMerge(const JoinAggArgs &a, uint64_t *acc, const uint64_t *t) {
#pragma unroll
	for (int q = 0; q < kJoinAggMax; q++) {
		if (q >= a.naggs) {
			continue;
		}
		uint64_t &x = acc[2 * q];
		uint64_t &y = acc[2 * q + 1];
		switch (a.kind[q]) {
		case JA_COUNT:
		case JA_SUM_I128:
			JaAdd128(x, y, t[2 * q], t[2 * q + 1]);
			break;
		case JA_SUM_F64:
			JaDdAdd(x, y, __longlong_as_double((long long)t[2 * q]), __longlong_as_double((long long)t[2 * q + 1]));
			break;
		case JA_MIN_I64:
			if ((long long)t[2 * q] < (long long)x) {
				x = t[2 * q];
			}
			break;
		case JA_MAX_I64:
			if ((long long)t[2 * q] > (long long)x) {
				x = t[2 * q];
			}
			break;
		case JA_MIN_F64:
			if (JaLess(__longlong_as_double((long long)t[2 * q]), __longlong_as_double((long long)x))) {
				x = t[2 * q];
			}
			break;
		case JA_MAX_F64:
			if (JaLess(__longlong_as_double((long long)x), __longlong_as_double((long long)t[2 * q]))) {
				x = t[2 * q];
			}
			break;
		default:
			break;
		}
	}
}
#endif

//! match[i] = index of the B key equal to A key i (both runs ascending in the ordered-key sense, unique), or
//! -1. vtype: the key type of both runs.
void MergeMatch(const void *akeys, uint64_t na, const void *bkeys, uint64_t nb, int vtype, long long *match,
                hipStream_t stream);
void PairCounts(const long long *match, const long long *ca, const long long *cb, uint64_t na, uint64_t *pairs,
                hipStream_t stream);
void ExclusiveScanU64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_device, Scratch &scratch,
                      hipStream_t stream);
void ExpandPairs(const uint64_t *offsets, uint64_t na, const long long *match, const long long *cb,
                 const long long *starta, const long long *startb, const uint32_t *perma, const uint32_t *permb,
                 uint64_t total, uint32_t *out_left, uint32_t *out_right, hipStream_t stream);
//! result_device: naggs x 2 words (kJoinAggMax x 2 allocated).
void JoinAggregate(const JoinAggArgs &args, uint64_t *result_device, Scratch &scratch, hipStream_t stream);

} // namespace vgpu
