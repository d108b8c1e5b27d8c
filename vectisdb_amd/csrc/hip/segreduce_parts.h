// segreduce_parts.h -- pieces of the segmented reduction shared by the row reduce (segreduce.hip) and the
// run reduce fused into the last radix pass (runreduce.hip): output descriptors and the group writer, the
// two-level scan of per-tile head counts, the fix-up of groups that span tiles, and the component
// signatures compiled ahead of time. Everything here has internal linkage, so each translation unit that
// includes it gets its own kernels.
#pragma once

#include "agg_state.h"
#include "common.h"
#include "kernels.h"

namespace vgpu {
namespace {

//! Tile offsets come from a two-level scan of per-tile head counts: kScanTiles tiles per first-level block.
constexpr int kScanBlock = 1024;
constexpr int kScanPerThread = 4;
constexpr uint64_t kScanTiles = (uint64_t)kScanBlock * kScanPerThread; // tiles per scan block

struct AggOutDesc {
	int naggs = 0;
	int count_comp = 0;
	int kind[16] = {};
	int comp[16] = {};
	int comp_kind[16] = {};
	int out_type[16] = {};
	void *out[16] = {};
};

struct KeyOutDesc {
	KeyPack pack;
	void *out[kMaxKeyCols] = {};
};

__device__ __forceinline__ double I128ToDouble(u64 lo, u64 hi) {
	bool neg = (long long)hi < 0;
	if (neg) {
		lo = ~lo + 1;
		hi = ~hi + (lo == 0 ? 1 : 0);
	}
	double v = (double)hi * 18446744073709551616.0 + (double)lo;
	return neg ? -v : v;
}

__device__ __forceinline__ void StoreTyped(void *out, u64 gid, int vtype, long long iv, double dv, bool is_float) {
	switch (vtype) {
	case VT_INT16:
		((short *)out)[gid] = (short)iv;
		break;
	case VT_INT32:
		((int *)out)[gid] = (int)iv;
		break;
	case VT_INT64:
		((long long *)out)[gid] = is_float ? (long long)dv : iv;
		break;
	case VT_BOOLEAN:
		((unsigned char *)out)[gid] = (unsigned char)(iv != 0);
		break;
	case VT_FLOAT32:
		((float *)out)[gid] = is_float ? (float)dv : (float)iv;
		break;
	default:
		((double *)out)[gid] = is_float ? dv : (double)iv;
		break;
	}
}

//! Register-array element select with a runtime index, unrolled so the state never leaves VGPRs.
template <typename S = DynSig>
__device__ __forceinline__ u64 Pick(const u64 (&arr)[kMaxComponents], int index) {
	u64 v = 0;
#pragma unroll
	for (int c = 0; c < S::n; c++) {
		if (c == index) {
			v = arr[c];
		}
	}
	return v;
}

template <typename S = DynSig>
__device__ __forceinline__ void WriteGroup(const AggOutDesc &o, const CompState &s, u64 gid) {
	const u64 count = Pick<S>(s.a, o.count_comp);
	for (int i = 0; i < o.naggs; i++) {
		const int c = o.comp[i];
		const u64 sa = Pick<S>(s.a, c);
		const u64 sb = Pick<S>(s.b, c);
		void *out = o.out[i];
		switch (o.kind[i]) {
		case AGG_COUNT_STAR:
		case AGG_COUNT:
			((long long *)out)[gid] = (long long)count;
			break;
		case AGG_SUM:
			if (o.comp_kind[i] == COMP_SUM_I128) {
				if (o.out_type[i] == VT_HUGEINT) {
					((u64 *)out)[2 * gid] = sa;
					((u64 *)out)[2 * gid + 1] = sb;
				} else if (o.out_type[i] == VT_INT64) {
					((long long *)out)[gid] = (long long)sa;
				} else {
					StoreTyped(out, gid, o.out_type[i], 0, I128ToDouble(sa, sb), true);
				}
			} else {
				StoreTyped(out, gid, o.out_type[i], 0, DdValue(AsDouble(sa), AsDouble(sb)), true);
			}
			break;
		case AGG_MIN:
		case AGG_MAX:
			if (o.comp_kind[i] == COMP_MIN_F64 || o.comp_kind[i] == COMP_MAX_F64) {
				StoreTyped(out, gid, o.out_type[i], 0, AsDouble(sa), true);
			} else {
				StoreTyped(out, gid, o.out_type[i], (long long)sa, 0.0, false);
			}
			break;
		case AGG_AVG: {
			double sum = o.comp_kind[i] == COMP_SUM_I128 ? I128ToDouble(sa, sb) : DdValue(AsDouble(sa), AsDouble(sb));
			StoreTyped(out, gid, o.out_type[i], 0, sum / (double)count, true);
			break;
		}
		default:
			break;
		}
	}
}

__device__ __forceinline__ void WriteKeys(const KeyOutDesc &ko, u64 key, u64 gid) {
	for (int c = 0; c < ko.pack.ncols; c++) {
		const int bits = ko.pack.bits[c];
		u64 field = bits == 0 ? 0ull : ((key >> ko.pack.shift[c]) & (bits >= 64 ? ~0ull : ((1ull << bits) - 1)));
		u64 raw = UnorderedKey(field + ko.pack.min[c], ko.pack.width[c], ko.pack.is_signed[c] != 0);
		switch (ko.pack.width[c]) {
		case 1:
			((uint8_t *)ko.out[c])[gid] = (uint8_t)raw;
			break;
		case 2:
			((uint16_t *)ko.out[c])[gid] = (uint16_t)raw;
			break;
		case 4:
			((uint32_t *)ko.out[c])[gid] = (uint32_t)raw;
			break;
		default:
			((u64 *)ko.out[c])[gid] = raw;
			break;
		}
	}
}



//! Loads ITEMS consecutive elements of type T at row `row0` as raw bits: 16-byte vector loads when all of
//! them are there and the address allows, element loads otherwise.
template <int ITEMS, typename T>
__device__ __forceinline__ void LoadRun(const T *src, uint64_t row0, uint64_t n, u64 (&out)[ITEMS]) {
	constexpr int kVecs = (int)(ITEMS * sizeof(T) / 16);
	if (kVecs > 0 && row0 + ITEMS <= n && (reinterpret_cast<uintptr_t>(src + row0) & 15) == 0) {
		T t[ITEMS];
		uint4 v[kVecs > 0 ? kVecs : 1];
#pragma unroll
		for (int j = 0; j < kVecs; j++) {
			v[j] = reinterpret_cast<const uint4 *>(src + row0)[j];
		}
		__builtin_memcpy(t, v, sizeof(t));
#pragma unroll
		for (int i = 0; i < ITEMS; i++) {
			out[i] = (u64)t[i];
		}
	} else {
#pragma unroll
		for (int i = 0; i < ITEMS; i++) {
			out[i] = row0 + i < n ? (u64)src[row0 + i] : 0ull;
		}
	}
}


//! Level 1: exclusive prefix of the head counts inside each run of kScanTiles tiles, and each run's total.
__global__ void __launch_bounds__(kScanBlock) TilePrefixKernel(const uint32_t *counts, uint64_t ntiles,
                                                               u64 *prefix, u64 *block_totals,
                                                               const uint64_t *n_dev, uint64_t tile_rows) {
	__shared__ u64 s_scan[kScanBlock / kWave + 1];
	if (n_dev != nullptr) {
		ntiles = (*n_dev + tile_rows - 1) / tile_rows;
		if ((uint64_t)blockIdx.x * kScanTiles >= ntiles) {
			return;
		}
	}
	const uint64_t t0 = (uint64_t)blockIdx.x * kScanTiles + (uint64_t)threadIdx.x * kScanPerThread;
	uint32_t c[kScanPerThread];
	u64 mine = 0;
#pragma unroll
	for (int i = 0; i < kScanPerThread; i++) {
		c[i] = t0 + i < ntiles ? counts[t0 + i] : 0u;
		mine += c[i];
	}
	u64 total;
	u64 run = BlockExclusiveScan<u64, kScanBlock>(mine, s_scan, &total);
#pragma unroll
	for (int i = 0; i < kScanPerThread; i++) {
		if (t0 + i < ntiles) {
			prefix[t0 + i] = run;
		}
		run += c[i];
	}
	if (threadIdx.x == 0) {
		block_totals[blockIdx.x] = total;
	}
}

//! Level 2 (one block): block_totals becomes each run's exclusive carry; *groups_out the total.
__global__ void __launch_bounds__(kScanBlock) TileCarryKernel(u64 *block_totals, uint64_t nblocks, u64 *groups_out,
                                                              const uint64_t *n_dev, uint64_t tile_rows) {
	__shared__ u64 s_scan[kScanBlock / kWave + 1];
	if (n_dev != nullptr) {
		nblocks = ((*n_dev + tile_rows - 1) / tile_rows + kScanTiles - 1) / kScanTiles;
	}
	u64 carry = 0;
	for (uint64_t base = 0; base < nblocks; base += kScanBlock) {
		const uint64_t i = base + threadIdx.x;
		const u64 v = i < nblocks ? block_totals[i] : 0ull;
		u64 total;
		const u64 excl = BlockExclusiveScan<u64, kScanBlock>(v, s_scan, &total);
		if (i < nblocks) {
			block_totals[i] = carry + excl;
		}
		carry += total;
	}
	if (threadIdx.x == 0) {
		*groups_out = carry;
	}
}


constexpr int kFixBlock = 256;
constexpr int kFixWaves = kFixBlock / kWave;

//! Level 1 of the fix-up. For each tile u of a 256-tile chunk: run[u] = lead[u] o lead[u+1] o ... up to
//! and including the first tile v >= u that starts a group (its lead is the part before that group), and
//! open[u] = no such tile inside the chunk (the run continues into the next chunk). A reverse segmented
//! scan: wave shuffles, then waves right to left through LDS. Tiles past the input act as a stop.
__global__ void __launch_bounds__(kFixBlock) ChunkRunKernel(CompDesc d, const uint32_t *tile_heads,
                                                            const u64 *lead_state, u64 *run_state,
                                                            uint8_t *run_open, uint64_t ntiles,
                                                            const uint64_t *n_dev, uint64_t tile_rows) {
	__shared__ u64 s_state[kFixWaves][2 * kMaxComponents];
	__shared__ int s_flag[kFixWaves];
	const int tid = threadIdx.x;
	const int lane = tid & (kWave - 1);
	const int wave = tid / kWave;
	if (n_dev != nullptr) {
		ntiles = (*n_dev + tile_rows - 1) / tile_rows;
		if ((uint64_t)blockIdx.x * kFixBlock >= ntiles) {
			return;
		}
	}
	const uint64_t u = (uint64_t)blockIdx.x * kFixBlock + tid;
	const bool in_range = u < ntiles;
	bool F = !in_range || tile_heads[u] > 0; // stop flag
	CompState X;
	if (in_range) {
		StateLoad(d, lead_state + u * 2 * d.ncomp, X);
	} else {
		StateIdentity(d, X);
	}
	for (int off = 1; off < kWave; off <<= 1) {
		CompState dn;
#pragma unroll
		for (int c = 0; c < kMaxComponents; c++) {
			if (c < d.ncomp) {
				dn.a[c] = __shfl_down(X.a[c], off);
				dn.b[c] = __shfl_down(X.b[c], off);
			}
		}
		bool fd = __shfl_down((int)F, off) != 0;
		if (lane + off < kWave) {
			if (!F) {
				StateCombine(d, X, dn);
			}
			F = F || fd;
		}
	}
	if (lane == 0) {
		StateStore(d, s_state[wave], X);
		s_flag[wave] = F ? 1 : 0;
	}
	__syncthreads();
	if (!F) {
		for (int w = wave + 1; w < kFixWaves; w++) {
			CompState t;
			StateLoad(d, s_state[w], t);
			StateCombine(d, X, t);
			if (s_flag[w]) {
				F = true;
				break;
			}
		}
	}
	if (in_range) {
		StateStore(d, run_state + u * 2 * d.ncomp, X);
		run_open[u] = F ? 0 : 1;
	}
}

//! Level 2: each open tail T folds run[T+1] and, while the run stays open, the first-tile runs of the
//! following chunks (a few steps even for a group spanning thousands of tiles), then writes the group.
__global__ void __launch_bounds__(kFixBlock) TailFinishKernel(CompDesc d, AggOutDesc o, const u64 *run_state,
                                                              const uint8_t *run_open, const u64 *tail_state,
                                                              const u64 *tail_gid, uint64_t ntiles,
                                                              const uint64_t *n_dev, uint64_t tile_rows) {
	if (n_dev != nullptr) {
		ntiles = (*n_dev + tile_rows - 1) / tile_rows;
	}
	for (uint64_t T = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; T < ntiles;
	     T += (uint64_t)gridDim.x * blockDim.x) {
		if (tail_gid[T] == ~0ull) {
			continue; // this tile's last group ended inside it
		}
		CompState acc;
		StateLoad(d, tail_state + T * 2 * d.ncomp, acc);
		uint64_t u = T + 1;
		while (u < ntiles) {
			CompState r;
			StateLoad(d, run_state + u * 2 * d.ncomp, r);
			StateCombine(d, acc, r);
			if (!run_open[u]) {
				break;
			}
			u = (u / kFixBlock + 1) * kFixBlock; // first tile of the next chunk
		}
		WriteGroup(o, acc, tail_gid[T]);
	}
}


//! Component signatures compiled ahead of time (COUNT plus the aggregates of the common queries); any other
//! mix runs the generic kernel.
using SigCountSumMinMaxDd = StaticSig<SigCode(COMP_COUNT), SigCode(COMP_SUM_I128, VT_INT64, 0),
                                      SigCode(COMP_MIN_I64, VT_INT64, 0), SigCode(COMP_MAX_I64, VT_INT64, 0),
                                      SigCode(COMP_SUM_DD, VT_FLOAT64, 1)>;
//! The same with the INT64 column carried as INT32 (the sort's payload narrowing).
using SigCountSumMinMaxDdN = StaticSig<SigCode(COMP_COUNT), SigCode(COMP_SUM_I128, VT_INT32, 0),
                                       SigCode(COMP_MIN_I64, VT_INT32, 0), SigCode(COMP_MAX_I64, VT_INT32, 0),
                                       SigCode(COMP_SUM_DD, VT_FLOAT64, 1)>;
using SigCountSumI32 = StaticSig<SigCode(COMP_COUNT), SigCode(COMP_SUM_I128, VT_INT32, 0)>;
using SigCountSumI64 = StaticSig<SigCode(COMP_COUNT), SigCode(COMP_SUM_I128, VT_INT64, 0)>;
using SigCountSumDd = StaticSig<SigCode(COMP_COUNT), SigCode(COMP_SUM_DD, VT_FLOAT64, 0)>;

template <typename SG>
bool SigMatches(const CompDesc &cd) {
	if (cd.ncomp != SG::n) {
		return false;
	}
	for (int c = 0; c < cd.ncomp; c++) {
		const int code = cd.kind[c] == COMP_COUNT ? SigCode(COMP_COUNT)
		                                          : SigCode(cd.kind[c], cd.in_type[c], cd.payload[c]);
		if (code != SG::code(c)) {
			return false;
		}
	}
	return true;
}


//! Kernel-side descriptors of a ReduceSpec: components, output aggregates and the key unpacking.
inline void BuildReduceDescs(const ReduceSpec &spec, CompDesc &cd, AggOutDesc &od, KeyOutDesc &ko) {
	cd = CompDesc {};
	cd.ncomp = spec.ncomp;
	for (int c = 0; c < spec.ncomp; c++) {
		cd.kind[c] = spec.comp[c].kind;
		cd.payload[c] = spec.comp[c].payload;
		cd.in_type[c] = spec.comp[c].in_type;
	}
	od = AggOutDesc {};
	od.naggs = spec.naggs;
	od.count_comp = 0;
	for (int c = 0; c < spec.ncomp; c++) {
		if (spec.comp[c].kind == COMP_COUNT) {
			od.count_comp = c;
			break;
		}
	}
	for (int i = 0; i < spec.naggs; i++) {
		od.kind[i] = spec.aggs[i].kind;
		od.comp[i] = spec.aggs[i].comp;
		od.comp_kind[i] = spec.comp[spec.aggs[i].comp].kind;
		od.out_type[i] = spec.aggs[i].out_type;
		od.out[i] = spec.aggs[i].out;
	}
	ko = KeyOutDesc {};
	ko.pack = spec.pack;
	for (int c = 0; c < spec.pack.ncols; c++) {
		ko.out[c] = spec.key_out[c];
	}
}

} // namespace
} // namespace vgpu
