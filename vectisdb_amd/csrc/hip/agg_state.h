// agg_state.h -- partial aggregate state shared by the segmented reduction and the join combiner.
//
// A group's state is kMaxComponents slots of two 64-bit words. Components (see kernels.h CompKind):
//   COUNT     a = rows
//   SUM_I128  (a, b) = two's-complement int128 (low, high); exact, so any combine order gives the same bits
//   SUM_DD    (a, b) = double-double (hi, lo); error-free TwoSum accumulation, rounded once at finalize, so
//             the result is the correctly rounded sum except under extreme cancellation
//   MIN/MAX   a = widened int64, or double bits with DuckDB's order (NaN greatest)
#pragma once

#include "common.h"
#include "kernels.h"

namespace vgpu {

struct CompState {
	u64 a[kMaxComponents];
	u64 b[kMaxComponents];
};

__device__ __forceinline__ double AsDouble(u64 bits) {
	return __longlong_as_double((long long)bits);
}
__device__ __forceinline__ u64 AsBits(double v) {
	return (u64)__double_as_longlong(v);
}

//! Error-free transformation: s + e == a + b exactly (round-to-nearest, no contraction).
__device__ __forceinline__ void TwoSum(double a, double b, double &s, double &e) {
	s = a + b;
	double bb = s - a;
	e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void DdAdd(double &hi, double &lo, double x) {
	double s, e;
	TwoSum(hi, x, s, e);
	e += lo;
	hi = s + e;
	lo = e - (hi - s);
}

__device__ __forceinline__ void DdAddDd(double &hi, double &lo, double hi2, double lo2) {
	double s, e;
	TwoSum(hi, hi2, s, e);
	e += lo + lo2;
	hi = s + e;
	lo = e - (hi - s);
}

__device__ __forceinline__ double DdValue(double hi, double lo) {
	if (!isfinite(hi)) {
		return hi;
	}
	return hi + lo;
}

__device__ __forceinline__ void I128AddI64(u64 &lo, u64 &hi, long long x) {
	u64 nlo = lo + (u64)x;
	u64 carry = nlo < lo ? 1ull : 0ull;
	hi = hi + (x < 0 ? ~0ull : 0ull) + carry;
	lo = nlo;
}

__device__ __forceinline__ void I128AddI128(u64 &lo, u64 &hi, u64 lo2, u64 hi2) {
	u64 nlo = lo + lo2;
	u64 carry = nlo < lo ? 1ull : 0ull;
	hi = hi + hi2 + carry;
	lo = nlo;
}

//! DuckDB orders NaN above every other double.
__device__ __forceinline__ bool DuckLess(double x, double y) {
	if (isnan(y)) {
		return !isnan(x);
	}
	return !isnan(x) && x < y;
}

struct CompDesc {
	int ncomp;
	int kind[kMaxComponents];
	int payload[kMaxComponents];
	int in_type[kMaxComponents];
};

//! Component signatures. DynSig reads each component's kind, input type and payload column from the
//! CompDesc at run time (any mix of up to kMaxComponents). StaticSig<K...> fixes them at compile time, one
//! code per component: kind | in_type << 4 | (payload + 1) << 8 (COUNT is just its kind), so once the
//! component loops unroll every switch folds away and only the arithmetic of those components remains.
struct DynSig {
	static constexpr bool kStatic = false;
	static constexpr int n = kMaxComponents;
	__host__ __device__ static constexpr int code(int) {
		return 0;
	}
	__host__ __device__ static constexpr bool has(int) {
		return true;
	}
};

template <int... K>
struct StaticSig {
	static constexpr bool kStatic = true;
	static constexpr int n = sizeof...(K);
	__host__ __device__ static constexpr int code(int c) {
		int i = 0, r = 0;
		((r = (i++ == c) ? K : r), ...);
		return r;
	}
	__host__ __device__ static constexpr bool has(int) {
		return true;
	}
};

//! The components of a static signature selected by MASK (bit c = component c): every state operation skips
//! the others, so a kernel can fold and scan one payload column's components at a time with only their
//! registers live. Component positions (and so the stored state layout) are those of SG.
template <typename SG, unsigned MASK>
struct MaskSig {
	static constexpr bool kStatic = true;
	static constexpr int n = SG::n;
	__host__ __device__ static constexpr int code(int c) {
		return SG::code(c);
	}
	__host__ __device__ static constexpr bool has(int c) {
		return ((MASK >> c) & 1u) != 0;
	}
};

constexpr int SigCode(int kind, int in_type = 0, int payload = -1) {
	return kind == COMP_COUNT ? COMP_COUNT : (kind | (in_type << 4) | ((payload + 1) << 8));
}

template <typename S>
__device__ __forceinline__ bool CHas(const CompDesc &d, int c) {
	if constexpr (S::kStatic) {
		return S::has(c);
	} else {
		return c < d.ncomp;
	}
}
template <typename S>
__device__ __forceinline__ int CKind(const CompDesc &d, int c) {
	if constexpr (S::kStatic) {
		return S::code(c) & 15;
	} else {
		return d.kind[c];
	}
}
template <typename S>
__device__ __forceinline__ int CType(const CompDesc &d, int c) {
	if constexpr (S::kStatic) {
		return (S::code(c) >> 4) & 15;
	} else {
		return d.in_type[c];
	}
}
template <typename S>
__device__ __forceinline__ int CPayload(const CompDesc &d, int c) {
	if constexpr (S::kStatic) {
		return (S::code(c) >> 8) - 1;
	} else {
		return d.payload[c];
	}
}
//! The component keeps a second word (SUM_I128 high, SUM_DD low).
template <typename S>
__device__ __forceinline__ bool CHasB(const CompDesc &d, int c) {
	const int k = CKind<S>(d, c);
	return k == COMP_SUM_I128 || k == COMP_SUM_DD;
}

template <typename S = DynSig>
__device__ __forceinline__ void StateIdentity(const CompDesc &d, CompState &s) {
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		s.a[c] = 0;
		s.b[c] = 0;
		if (CHas<S>(d, c)) {
			switch (CKind<S>(d, c)) {
			case COMP_MIN_I64:
				s.a[c] = (u64)0x7fffffffffffffffll;
				break;
			case COMP_MAX_I64:
				s.a[c] = (u64)0x8000000000000000ull;
				break;
			case COMP_MIN_F64:
				s.a[c] = AsBits(__builtin_nan(""));
				break;
			case COMP_MAX_F64:
				s.a[c] = AsBits(-__builtin_inf());
				break;
			default:
				break;
			}
		}
	}
}

//! Value of a payload element as int64 (integers, sign-extended; BOOLEAN 0/1) -- raw bits in `raw`.
__device__ __forceinline__ long long RawAsI64(u64 raw, int vtype) {
	switch (vtype) {
	case VT_INT16:
		return (long long)(short)(unsigned short)raw;
	case VT_INT32:
		return (long long)(int)(unsigned)raw;
	case VT_BOOLEAN:
		return (long long)((raw & 0xff) != 0);
	default:
		return (long long)raw;
	}
}

__device__ __forceinline__ double RawAsF64(u64 raw, int vtype) {
	if (vtype == VT_FLOAT32) {
		return (double)__uint_as_float((unsigned)raw);
	}
	if (vtype == VT_FLOAT64) {
		return AsDouble(raw);
	}
	return (double)RawAsI64(raw, vtype);
}

//! Folds one row into the state. `vals[p]` holds the raw bits of payload column p at this row.
template <int NP, typename S = DynSig>
__device__ __forceinline__ void StateAddRow(const CompDesc &d, CompState &s, const u64 (&vals)[NP]) {
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		if (CHas<S>(d, c)) {
			const int p = CPayload<S>(d, c);
			u64 raw = 0;
#pragma unroll
			for (int q = 0; q < NP; q++) {
				if (q == p) {
					raw = vals[q];
				}
			}
			const int t = CType<S>(d, c);
			switch (CKind<S>(d, c)) {
			case COMP_COUNT:
				s.a[c] += 1;
				break;
			case COMP_SUM_I128:
				I128AddI64(s.a[c], s.b[c], RawAsI64(raw, t));
				break;
			case COMP_SUM_DD: {
				double hi = AsDouble(s.a[c]), lo = AsDouble(s.b[c]);
				DdAdd(hi, lo, RawAsF64(raw, t));
				s.a[c] = AsBits(hi);
				s.b[c] = AsBits(lo);
				break;
			}
			case COMP_MIN_I64: {
				long long v = RawAsI64(raw, t);
				if (v < (long long)s.a[c]) {
					s.a[c] = (u64)v;
				}
				break;
			}
			case COMP_MAX_I64: {
				long long v = RawAsI64(raw, t);
				if (v > (long long)s.a[c]) {
					s.a[c] = (u64)v;
				}
				break;
			}
			case COMP_MIN_F64: {
				double v = RawAsF64(raw, t);
				if (DuckLess(v, AsDouble(s.a[c]))) {
					s.a[c] = AsBits(v);
				}
				break;
			}
			case COMP_MAX_F64: {
				double v = RawAsF64(raw, t);
				if (DuckLess(AsDouble(s.a[c]), v)) {
					s.a[c] = AsBits(v);
				}
				break;
			}
			default:
				break;
			}
		}
	}
}

//! s = s (earlier rows) combined with t (later rows).
template <typename S = DynSig>
__device__ __forceinline__ void StateCombine(const CompDesc &d, CompState &s, const CompState &t) {
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		if (CHas<S>(d, c)) {
			switch (CKind<S>(d, c)) {
			case COMP_COUNT:
				s.a[c] += t.a[c];
				break;
			case COMP_SUM_I128:
				I128AddI128(s.a[c], s.b[c], t.a[c], t.b[c]);
				break;
			case COMP_SUM_DD: {
				double hi = AsDouble(s.a[c]), lo = AsDouble(s.b[c]);
				DdAddDd(hi, lo, AsDouble(t.a[c]), AsDouble(t.b[c]));
				s.a[c] = AsBits(hi);
				s.b[c] = AsBits(lo);
				break;
			}
			case COMP_MIN_I64:
				if ((long long)t.a[c] < (long long)s.a[c]) {
					s.a[c] = t.a[c];
				}
				break;
			case COMP_MAX_I64:
				if ((long long)t.a[c] > (long long)s.a[c]) {
					s.a[c] = t.a[c];
				}
				break;
			case COMP_MIN_F64:
				if (DuckLess(AsDouble(t.a[c]), AsDouble(s.a[c]))) {
					s.a[c] = t.a[c];
				}
				break;
			case COMP_MAX_F64:
				if (DuckLess(AsDouble(s.a[c]), AsDouble(t.a[c]))) {
					s.a[c] = t.a[c];
				}
				break;
			default:
				break;
			}
		}
	}
}

template <typename S = DynSig>
__device__ __forceinline__ void StateShflUp(const CompDesc &d, CompState &dst, const CompState &src, int delta) {
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		if (CHas<S>(d, c)) {
			dst.a[c] = __shfl_up(src.a[c], delta);
			if (!S::kStatic || CHasB<S>(d, c)) {
				dst.b[c] = __shfl_up(src.b[c], delta);
			} else {
				dst.b[c] = 0;
			}
		}
	}
}

template <typename S = DynSig>
__device__ __forceinline__ void StateShflXor(const CompDesc &d, CompState &dst, const CompState &src, int mask) {
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		if (CHas<S>(d, c)) {
			dst.a[c] = __shfl_xor(src.a[c], mask);
			if (!S::kStatic || CHasB<S>(d, c)) {
				dst.b[c] = __shfl_xor(src.b[c], mask);
			} else {
				dst.b[c] = 0;
			}
		}
	}
}

template <typename S = DynSig>
__device__ __forceinline__ void StateStore(const CompDesc &d, u64 *dst, const CompState &s) {
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		if (CHas<S>(d, c)) {
			dst[2 * c] = s.a[c];
			dst[2 * c + 1] = s.b[c];
		}
	}
}

template <typename S = DynSig>
__device__ __forceinline__ void StateLoad(const CompDesc &d, const u64 *src, CompState &s) {
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		if (CHas<S>(d, c)) {
			s.a[c] = src[2 * c];
			s.b[c] = src[2 * c + 1];
		} else {
			s.a[c] = 0;
			s.b[c] = 0;
		}
	}
}

} // namespace vgpu
