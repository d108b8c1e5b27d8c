// segreduce.hip -- one-pass segmented multi-aggregate reduction over a key-sorted run (gfx950).
//
// Replaces the reference's head flags + inclusive_scan group ids + group-count readback + one
// `thrust::reduce_by_key` (preceded by a `GatherColumn`) PER AGGREGATE
// (cuda_engine/src/operators/gpu_groupby.cu:380-576). Here every aggregate of the query is folded in the
// same pass over the sorted keys and payload:
//   * each thread reduces 8 consecutive rows serially; groups that start and end inside the thread are
//     finished right there;
//   * open segments are stitched across the 256 threads with a wave64 segmented scan (shuffles) plus one
//     LDS step across the 4 waves;
//   * a group's output index is its head's global rank, which each tile learns by decoupled look-back on
//     per-tile head counts (one 64-bit atomic word per tile);
//   * a group that continues past its tile leaves a tail partial; a two-level fix-up (a reverse segmented
//     scan of tile leads per 256-tile chunk, then one fold per open tail plus one per spanned chunk)
//     finishes it, so even the head group of a Zipf key (thousands of tiles) costs a few dependent loads.
// SUM over integers accumulates in exact int128 (DuckDB's HUGEINT), SUM/AVG over doubles in double-double
// so the result is rounded once.
#include "agg_state.h"
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace vgpu {
namespace {

constexpr int kRedBlock = 256;
constexpr int kRedItems = 8;
constexpr int kRedTile = kRedBlock * kRedItems;
constexpr int kRedWaves = kRedBlock / kWave;
//! Tile offsets come from a two-level scan of per-tile head counts: kScanTiles tiles per first-level block.
constexpr int kScanBlock = 1024;
constexpr int kScanPerThread = 4;
constexpr uint64_t kScanTiles = (uint64_t)kScanBlock * kScanPerThread; // tiles per scan block

struct PayloadPtrs {
	int n = 0;
	const void *ptr[kMaxPayload] = {};
	int width[kMaxPayload] = {};
};

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

//! Copies `nbytes` from global `src` into LDS `dst` with the whole block: 16-byte vector loads when both
//! ends allow it, 4-byte otherwise, bytes for any ragged tail. Consecutive threads read consecutive
//! vectors, so the copy is one coalesced sweep of the tile.
template <int MAXV>
__device__ __forceinline__ void StageToLds(uint8_t *dst, const uint8_t *src, uint32_t nbytes) {
	uint32_t done = 0;
	if ((((uintptr_t)src) & 15) == 0) {
		// All of this thread's vectors are loaded before any is stored, so MAXV loads are in flight at once.
		const uint32_t vecs = nbytes / 16;
		uint4 v[MAXV];
#pragma unroll
		for (int j = 0; j < MAXV; j++) {
			uint32_t i = threadIdx.x + j * kRedBlock;
			if (i < vecs) {
				v[j] = reinterpret_cast<const uint4 *>(src)[i];
			}
		}
#pragma unroll
		for (int j = 0; j < MAXV; j++) {
			uint32_t i = threadIdx.x + j * kRedBlock;
			if (i < vecs) {
				reinterpret_cast<uint4 *>(dst)[i] = v[j];
			}
		}
		done = vecs * 16;
	} else if ((((uintptr_t)src) & 3) == 0) {
		const uint32_t words = nbytes / 4;
		for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) {
			reinterpret_cast<uint32_t *>(dst)[i] = reinterpret_cast<const uint32_t *>(src)[i];
		}
		done = words * 4;
	}
	for (uint32_t i = done + threadIdx.x; i < nbytes; i += blockDim.x) {
		dst[i] = src[i];
	}
}

__device__ __forceinline__ u64 LdsRawBits(const uint8_t *base, uint32_t row, int width) {
	switch (width) {
	case 1:
		return base[row];
	case 2:
		return reinterpret_cast<const uint16_t *>(base)[row];
	case 4:
		return reinterpret_cast<const uint32_t *>(base)[row];
	default:
		return reinterpret_cast<const u64 *>(base)[row];
	}
}

constexpr int kSlotStride = kRedBlock + 8; // 264 = 8 (mod 64): conflict-free transposed staging

//! Row r of the tile -> dst[(r % 8) * kSlotStride + r / 8]. Each thread loads rows tid + 256 j (coalesced
//! across the wave), all loads first, then the stores.
template <typename T, typename D>
__device__ __forceinline__ void StageTransposed(D *dst, const T *src, uint32_t count) {
	T v[kRedItems];
#pragma unroll
	for (int j = 0; j < kRedItems; j++) {
		uint32_t r = threadIdx.x + j * kRedBlock;
		v[j] = r < count ? src[r] : (T)0;
	}
#pragma unroll
	for (int j = 0; j < kRedItems; j++) {
		uint32_t r = threadIdx.x + j * kRedBlock;
		if (r < count) {
			dst[(r % kRedItems) * kSlotStride + r / kRedItems] = (D)v[j];
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

//! Rows per thread of the reduce: 8 in the generic kernel (LDS staging), kRedItemsDirect in the
//! specialized ones (register loads, so the cross-thread scan is paid once per more rows).
constexpr int kRedItemsDirect = 8;

template <typename SG>
constexpr int RedItems() {
	return SG::kStatic ? kRedItemsDirect : kRedItems;
}

//! Reduce tile: 256 threads x 8 consecutive rows, each thread folding its rows serially. The generic kernel
//! stages keys and payload columns into LDS with one coalesced sweep each (transposed, conflict-free); the
//! kernels specialized for one component signature (DIRECT) load each thread's 8 rows straight into
//! registers with 16-byte vector loads instead, which frees the 43 KB of LDS that capped occupancy.
template <typename KeyT, int NP, typename SG>
__global__ void __launch_bounds__(kRedBlock)
    SegReduceKernel(const KeyT *keys, uint64_t n, PayloadPtrs pay, CompDesc d, AggOutDesc o, KeyOutDesc ko,
                    const u64 *tile_prefix, const u64 *tile_carry, u64 *lead_state, u64 *tail_state, u64 *tail_gid,
                    uint64_t ntiles, const uint64_t *n_dev) {
	__shared__ u64 s_scan[kRedWaves + 1];
	__shared__ u64 s_wave_state[kRedWaves][2 * kMaxComponents];
	__shared__ int s_wave_flag[kRedWaves];
	constexpr bool DIRECT = SG::kStatic;
	constexpr int ITEMS = RedItems<SG>();
	constexpr int TILE = kRedBlock * ITEMS;
	constexpr int kStage = DIRECT ? 1 : kRedItems * kSlotStride;
	__shared__ KeyT s_keys[kStage];
	__shared__ KeyT s_prev;
	__shared__ u64 s_pay[NP > 0 ? NP : 1][kStage]; // raw bits, zero-extended
	if (n_dev != nullptr) {
		n = *n_dev;
		if ((uint64_t)blockIdx.x * TILE >= n) {
			return;
		}
	}

	const int tid = threadIdx.x;
	const int lane = tid & (kWave - 1);
	const int wave = tid / kWave;
	// Static tiles: the group ids before this tile come from the head-count scan, so no tile waits on
	// another and no tile takes a ticket (one same-address atomic per tile serializes at ~12 ns each).
	const uint32_t tile = blockIdx.x;
	const uint64_t tile_base = (uint64_t)tile * TILE;
	const uint32_t tile_count = (uint32_t)((n - tile_base) < (uint64_t)TILE ? (n - tile_base) : TILE);

	// DIRECT: this thread's 8 rows in registers (raw bits) and the key of the row before them.
	u64 kreg[ITEMS];
	u64 preg[NP > 0 ? NP : 1][ITEMS];
	KeyT kprev = (KeyT)0;
	if constexpr (DIRECT) {
		const uint64_t row0 = tile_base + (uint64_t)tid * ITEMS;
		if (keys != nullptr) {
			LoadRun<ITEMS, KeyT>(keys, row0, n, kreg);
			kprev = row0 > 0 && row0 <= n ? keys[row0 - 1] : (KeyT)0;
		}
#pragma unroll
		for (int p = 0; p < NP; p++) {
			if (p < pay.n) {
				switch (pay.width[p]) {
				case 8:
					LoadRun<ITEMS, u64>(reinterpret_cast<const u64 *>(pay.ptr[p]), row0, n, preg[p]);
					break;
				case 4:
					LoadRun<ITEMS, uint32_t>(reinterpret_cast<const uint32_t *>(pay.ptr[p]), row0, n, preg[p]);
					break;
				case 2:
					LoadRun<ITEMS, uint16_t>(reinterpret_cast<const uint16_t *>(pay.ptr[p]), row0, n, preg[p]);
					break;
				default:
					LoadRun<ITEMS, uint8_t>(reinterpret_cast<const uint8_t *>(pay.ptr[p]), row0, n, preg[p]);
					break;
				}
			}
		}
	}
	// Stage, transposed: row r of the tile goes to slot (r % 8) * kSlotStride + r / 8, so a thread's 8
	// consecutive rows are 8 slots one stride apart and a wave reads 64 consecutive slots per step; the
	// stride (264 = 8 mod 64) keeps the staging writes conflict-free too. Loads are coalesced per wave and
	// all issued before the first LDS store. s_prev = key before the tile.
	if (!DIRECT && keys != nullptr) {
		StageTransposed<KeyT>(s_keys, keys + tile_base, tile_count);
		if (tid == 0) {
			s_prev = tile_base > 0 ? keys[tile_base - 1] : (KeyT)0;
		}
	}
#pragma unroll
	for (int p = 0; p < NP; p++) {
		if (!DIRECT && p < pay.n) {
			const int w = pay.width[p];
			switch (w) {
			case 8:
				StageTransposed<u64>(s_pay[p], reinterpret_cast<const u64 *>(pay.ptr[p]) + tile_base, tile_count);
				break;
			case 4:
				StageTransposed<uint32_t>(s_pay[p], reinterpret_cast<const uint32_t *>(pay.ptr[p]) + tile_base,
				                          tile_count);
				break;
			case 2:
				StageTransposed<uint16_t>(s_pay[p], reinterpret_cast<const uint16_t *>(pay.ptr[p]) + tile_base,
				                          tile_count);
				break;
			default:
				StageTransposed<uint8_t>(s_pay[p], reinterpret_cast<const uint8_t *>(pay.ptr[p]) + tile_base,
				                         tile_count);
				break;
			}
		}
	}
	__syncthreads();

	const uint32_t local0 = (uint32_t)tid * ITEMS;
	const uint64_t base = tile_base + local0;
	KeyT k[ITEMS];
	bool head[ITEMS];
	u64 hc = 0;
	KeyT before = keys == nullptr ? (KeyT)0
	                              : (DIRECT ? kprev
	                                        : (tid == 0 ? s_prev : s_keys[(kRedItems - 1) * kSlotStride + tid - 1]));
#pragma unroll
	for (int i = 0; i < ITEMS; i++) {
		uint32_t lr = local0 + i;
		if constexpr (DIRECT) {
			k[i] = (keys != nullptr && lr < tile_count) ? (KeyT)kreg[i] : (KeyT)0;
		} else {
			k[i] = (keys != nullptr && lr < tile_count) ? s_keys[i * kSlotStride + tid] : (KeyT)0;
		}
		uint64_t row = base + i;
		head[i] = lr < tile_count && (row == 0 || k[i] != before);
		before = k[i];
		hc += head[i] ? 1 : 0;
	}
	u64 H;
	u64 hb_excl = BlockExclusiveScan<u64, kRedBlock>(hc, s_scan, &H);
	const u64 hb_global = tile_prefix[tile] + tile_carry[tile / kScanTiles];

	// Serial pass over this thread's rows.
	CompState S, lead;
	StateIdentity<SG>(d, S);
	StateIdentity<SG>(d, lead);
	bool seen = false;
	u64 gid = hb_global + hb_excl;
	u64 cur_gid = 0;
#pragma unroll
	for (int i = 0; i < ITEMS; i++) {
		uint32_t lr = local0 + i;
		if (lr < tile_count) {
			u64 vals[NP > 0 ? NP : 1];
#pragma unroll
			for (int p = 0; p < (NP > 0 ? NP : 1); p++) {
				if constexpr (DIRECT) {
					vals[p] = (p < pay.n) ? preg[p][i] : 0ull;
				} else {
					vals[p] = (p < pay.n) ? s_pay[p][i * kSlotStride + tid] : 0ull;
				}
			}
			if (head[i]) {
				if (seen) {
					WriteGroup<SG>(o, S, cur_gid);
				} else {
					lead = S;
					seen = true;
				}
				cur_gid = gid++;
				WriteKeys(ko, (u64)k[i], cur_gid);
				StateIdentity<SG>(d, S);
			}
			StateAddRow<(NP > 0 ? NP : 1), SG>(d, S, vals);
		}
	}
	if (!seen) {
		lead = S;
	}

	// Segmented inclusive scan of (seen, S) across the wave: X covers from the latest head at or before this
	// thread's last row.
	CompState X = S;
	bool F = seen;
	for (int off = 1; off < kWave; off <<= 1) {
		CompState up;
		StateShflUp<SG>(d, up, X, off);
		bool fup = __shfl_up((int)F, off) != 0;
		if (lane >= off) {
			if (!F) {
				StateCombine<SG>(d, up, X);
				X = up;
			}
			F = F || fup;
		}
	}
	if (lane == kWave - 1) {
		StateStore<SG>(d, s_wave_state[wave], X);
		s_wave_flag[wave] = F ? 1 : 0;
	}
	__syncthreads();
	CompState P;
	StateIdentity<SG>(d, P);
	for (int w = 0; w < wave; w++) {
		CompState t;
		StateLoad<SG>(d, s_wave_state[w], t);
		if (s_wave_flag[w]) {
			P = t;
		} else {
			StateCombine<SG>(d, P, t);
		}
	}
	if (!F) {
		CompState t = P;
		StateCombine<SG>(d, t, X);
		X = t;
	}
	CompState E;
	StateShflUp<SG>(d, E, X, 1);
	if (lane == 0) {
		E = P;
	}

	if (seen) {
		CompState V = E;
		StateCombine<SG>(d, V, lead);
		if (hb_excl > 0) {
			WriteGroup<SG>(o, V, hb_global + hb_excl - 1);
		} else {
			StateStore<SG>(d, lead_state + (uint64_t)tile * 2 * d.ncomp, V);
		}
	}
	if (tid == kRedBlock - 1) {
		uint64_t tile_end = tile_base + TILE;
		if (H > 0) {
			u64 gid_last = hb_global + H - 1;
			bool ends = tile_end >= n;
			if (!ends && keys != nullptr) {
				// keys == nullptr: one global group, which only ends with the input
				ends = keys[tile_end] != (DIRECT ? (KeyT)kreg[ITEMS - 1]
				                                  : s_keys[(kRedItems - 1) * kSlotStride + kRedBlock - 1]);
			}
			if (ends) {
				WriteGroup<SG>(o, X, gid_last);
				tail_gid[tile] = ~0ull;
			} else {
				StateStore<SG>(d, tail_state + (uint64_t)tile * 2 * d.ncomp, X);
				tail_gid[tile] = gid_last; // an open tail: TailFinishKernel completes the group
			}
		} else {
			StateStore<SG>(d, lead_state + (uint64_t)tile * 2 * d.ncomp, X);
			tail_gid[tile] = ~0ull;
		}
	}
	(void)ntiles;
	(void)wave;
	(void)lane;
}

//! Group heads per reduce tile (the same rows and the same head test as SegReduceKernel: row r starts a
//! group when r == 0 or its key differs from row r - 1's; without keys only row 0 does). One wave per
//! tile, each lane 32 consecutive keys read with vector loads; no barriers.
constexpr int kHeadLaneRows = 32;
template <typename KeyT>
__global__ void __launch_bounds__(256) HeadCountKernel(const KeyT *keys, uint64_t n, uint64_t tile_rows,
                                                      uint64_t ntiles, uint32_t *counts, const uint64_t *n_dev) {
	if (n_dev != nullptr) { // row count written by an earlier kernel; n was its bound
		n = *n_dev;
		ntiles = (n + tile_rows - 1) / tile_rows;
	}
	const uint64_t tile = (uint64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
	if (tile >= ntiles) {
		return;
	}
	const int lane = LaneId();
	u64 hc = 0;
	if (keys == nullptr) {
		hc = (tile == 0 && lane == 0 && n > 0) ? 1 : 0;
	} else {
		// tile_rows = 64 lanes x kHeadLaneRows rows
		const uint64_t row0 = tile * tile_rows + (uint64_t)lane * kHeadLaneRows;
		if (row0 < n) {
			u64 kreg[kHeadLaneRows];
			LoadRun<kHeadLaneRows, KeyT>(keys, row0, n, kreg);
			KeyT before = row0 > 0 ? keys[row0 - 1] : (KeyT)0;
#pragma unroll
			for (int i = 0; i < kHeadLaneRows; i++) {
				const KeyT k = (KeyT)kreg[i];
				hc += (row0 + i < n && (row0 + i == 0 || k != before)) ? 1 : 0;
				before = k;
			}
		}
	}
	for (int off = kWave / 2; off > 0; off >>= 1) {
		hc += __shfl_xor(hc, off);
	}
	if (lane == 0) {
		counts[tile] = (uint32_t)hc;
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

//! Count-only reduce (every aggregate is a COUNT): a run-length encoding of the sorted keys. One wave per
//! 2,048-row tile (the head-count tiles), each lane 32 consecutive keys in registers (vector loads; the
//! key before them is one extra load that hits the neighbour's cache line); the tile's first group id comes
//! from the head-count scan, so no tile takes a ticket or waits on another. Each group's key and first row
//! are written; RleCountKernel turns first rows into counts.
template <typename KeyT>
__global__ void __launch_bounds__(256) RleHeadsKernel(const KeyT *keys, uint64_t n, KeyOutDesc ko, u64 *starts,
                                                     const u64 *tile_prefix, const u64 *tile_carry,
                                                     uint64_t ntiles, const uint64_t *n_dev) {
	if (n_dev != nullptr) {
		n = *n_dev;
		ntiles = (n + (uint64_t)kWave * kHeadLaneRows - 1) / ((uint64_t)kWave * kHeadLaneRows);
	}
	const uint64_t tile = (uint64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
	if (tile >= ntiles) {
		return;
	}
	const int lane = LaneId();
	const uint64_t row0 = tile * ((uint64_t)kWave * kHeadLaneRows) + (uint64_t)lane * kHeadLaneRows;
	KeyT k[kHeadLaneRows];
	if (row0 + kHeadLaneRows <= n && (reinterpret_cast<uintptr_t>(keys) & 15) == 0) {
		// 32 consecutive keys: 128 B (u32) or 256 B (u64), 16-byte aligned (row0 is a multiple of 32)
		constexpr int kVecs = (int)(kHeadLaneRows * sizeof(KeyT) / 16);
		const uint4 *src = reinterpret_cast<const uint4 *>(keys + row0);
		uint4 v[kVecs];
#pragma unroll
		for (int j = 0; j < kVecs; j++) {
			v[j] = src[j];
		}
		__builtin_memcpy(k, v, sizeof(k));
	} else {
#pragma unroll
		for (int i = 0; i < kHeadLaneRows; i++) {
			k[i] = row0 + i < n ? keys[row0 + i] : (KeyT)0;
		}
	}
	KeyT before = row0 > 0 && row0 <= n ? keys[row0 - 1] : (KeyT)0;
	uint32_t heads = 0; // bit i = row0 + i starts a group
#pragma unroll
	for (int i = 0; i < kHeadLaneRows; i++) {
		const bool h = row0 + i < n && (row0 + i == 0 || k[i] != before);
		before = k[i];
		heads |= h ? (1u << i) : 0u;
	}
	const u64 hc = (u64)__popc(heads);
	u64 incl = hc;
	for (int off = 1; off < kWave; off <<= 1) {
		const u64 up = __shfl_up(incl, off);
		if (lane >= off) {
			incl += up;
		}
	}
	u64 gid = tile_prefix[tile] + tile_carry[tile / kScanTiles] + incl - hc;
#pragma unroll
	for (int i = 0; i < kHeadLaneRows; i++) {
		if ((heads >> i) & 1u) {
			WriteKeys(ko, (u64)k[i], gid);
			starts[gid] = row0 + i;
			gid++;
		}
	}
}

__global__ void RleCountKernel(const u64 *starts, const u64 *groups, uint64_t n, AggOutDesc o,
                               const uint64_t *n_dev) {
	if (n_dev != nullptr) {
		n = *n_dev;
	}
	const u64 g_total = *groups;
	const u64 stride = (u64)gridDim.x * blockDim.x;
	for (u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x; g < g_total; g += stride) {
		const u64 end = g + 1 < g_total ? starts[g + 1] : n;
		const long long c = (long long)(end - starts[g]);
		for (int i = 0; i < o.naggs; i++) {
			((long long *)o.out[i])[g] = c;
		}
	}
}

template <typename KeyT>
void LaunchRle(const SortedRun &run, uint64_t n, uint64_t capacity, const AggOutDesc &od, const KeyOutDesc &ko,
               u64 *groups_device, Scratch &scratch, hipStream_t stream, const uint64_t *n_dev) {
	constexpr uint64_t kTile = (uint64_t)kWave * kHeadLaneRows;
	const uint64_t ntiles = (n + kTile - 1) / kTile;
	const uint64_t nscan = (ntiles + kScanTiles - 1) / kScanTiles;
	auto *tile_heads = static_cast<uint32_t *>(scratch.Alloc(std::max<uint64_t>(ntiles, 1) * sizeof(uint32_t)));
	auto *tile_prefix = static_cast<u64 *>(scratch.Alloc(std::max<uint64_t>(ntiles, 1) * sizeof(u64)));
	auto *tile_carry = static_cast<u64 *>(scratch.Alloc(std::max<uint64_t>(nscan, 1) * sizeof(u64)));
	auto *starts = static_cast<u64 *>(scratch.Alloc(std::max<uint64_t>(capacity, 1) * sizeof(u64)));
	if (n == 0) {
		FillU64(groups_device, 0ull, 1, stream);
		return;
	}
	{
		ProfileScope prof("seg_count", stream, (double)n * sizeof(KeyT));
		hipLaunchKernelGGL((HeadCountKernel<KeyT>), dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, stream,
		                   (const KeyT *)run.keys, n, kTile, ntiles, tile_heads, n_dev);
		VGPU_HIP(hipGetLastError());
		hipLaunchKernelGGL(TilePrefixKernel, dim3((unsigned)nscan), dim3(kScanBlock), 0, stream, tile_heads, ntiles,
		                   tile_prefix, tile_carry, n_dev, kTile);
		VGPU_HIP(hipGetLastError());
		hipLaunchKernelGGL(TileCarryKernel, dim3(1), dim3(kScanBlock), 0, stream, tile_carry, nscan, groups_device,
		                   n_dev, kTile);
		VGPU_HIP(hipGetLastError());
		hipLaunchKernelGGL((RleHeadsKernel<KeyT>), dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, stream,
		                   (const KeyT *)run.keys, n, ko, starts, tile_prefix, tile_carry, ntiles, n_dev);
		VGPU_HIP(hipGetLastError());
	}
	const unsigned grid = (unsigned)std::min<uint64_t>((std::max<uint64_t>(capacity, 1) + 255) / 256, 8192);
	hipLaunchKernelGGL(RleCountKernel, dim3(grid), dim3(256), 0, stream, starts, groups_device, n, od, n_dev);
	VGPU_HIP(hipGetLastError());
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

template <typename KeyT, int NP, typename SG = DynSig>
void LaunchReduce(const SortedRun &run, uint64_t n, const PayloadPtrs &pay, const CompDesc &cd,
                  const AggOutDesc &od, const KeyOutDesc &ko, u64 *groups_device, Scratch &scratch,
                  hipStream_t stream, const uint64_t *n_dev) {
	constexpr int kItems = RedItems<SG>();
	constexpr uint64_t kTile = (uint64_t)kRedBlock * kItems;
	const uint64_t ntiles = (n + kTile - 1) / kTile;
	const size_t state_words = 2 * (size_t)std::max(cd.ncomp, 1);
	auto *tile_heads = static_cast<uint32_t *>(scratch.Alloc(ntiles * sizeof(uint32_t)));
	auto *tile_prefix = static_cast<u64 *>(scratch.Alloc(ntiles * sizeof(u64)));
	auto *lead_state = static_cast<u64 *>(scratch.Alloc(ntiles * state_words * sizeof(u64)));
	auto *tail_state = static_cast<u64 *>(scratch.Alloc(ntiles * state_words * sizeof(u64)));
	auto *tail_gid = static_cast<u64 *>(scratch.Alloc(ntiles * sizeof(u64)));
	const uint64_t nscan = (ntiles + kScanTiles - 1) / kScanTiles;
	auto *tile_carry = static_cast<u64 *>(scratch.Alloc(std::max<uint64_t>(nscan, 1) * sizeof(u64)));
	{
		static_assert(kTile == (uint64_t)kWave * kHeadLaneRows, "one wave per reduce tile");
		ProfileScope prof("seg_heads", stream, (double)n * (run.keys ? (run.keys64 ? 8.0 : 4.0) : 0.0));
		hipLaunchKernelGGL((HeadCountKernel<KeyT>), dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, stream,
		                   (const KeyT *)run.keys, n, kTile, ntiles, tile_heads, n_dev);
		VGPU_HIP(hipGetLastError());
		hipLaunchKernelGGL(TilePrefixKernel, dim3((unsigned)nscan), dim3(kScanBlock), 0, stream, tile_heads, ntiles,
		                   tile_prefix, tile_carry, n_dev, kTile);
		VGPU_HIP(hipGetLastError());
		hipLaunchKernelGGL(TileCarryKernel, dim3(1), dim3(kScanBlock), 0, stream, tile_carry, nscan, groups_device,
		                   n_dev, kTile);
		VGPU_HIP(hipGetLastError());
	}
	double bytes = (double)n * (run.keys ? (run.keys64 ? 8.0 : 4.0) : 0.0);
	for (int p = 0; p < pay.n; p++) {
		bytes += (double)n * pay.width[p];
	}
	{
		ProfileScope prof("seg_reduce", stream, bytes);
		hipLaunchKernelGGL((SegReduceKernel<KeyT, NP, SG>), dim3((unsigned)ntiles), dim3(kRedBlock), 0, stream,
		                   (const KeyT *)run.keys, n, pay, cd, od, ko, tile_prefix, tile_carry, lead_state, tail_state,
		                   tail_gid, ntiles, n_dev);
		VGPU_HIP(hipGetLastError());
	}
	{
		ProfileScope prof("seg_fixup", stream, 0.0);
		const uint64_t nchunks = (ntiles + kFixBlock - 1) / kFixBlock;
		auto *run_state = static_cast<u64 *>(scratch.Alloc(ntiles * state_words * sizeof(u64)));
		auto *run_open = static_cast<uint8_t *>(scratch.Alloc(ntiles));
		hipLaunchKernelGGL(ChunkRunKernel, dim3((unsigned)nchunks), dim3(kFixBlock), 0, stream, cd, tile_heads,
		                   lead_state, run_state, run_open, ntiles, n_dev, kTile);
		VGPU_HIP(hipGetLastError());
		hipLaunchKernelGGL(TailFinishKernel, dim3((unsigned)nchunks), dim3(kFixBlock), 0, stream, cd, od, run_state,
		                   run_open, tail_state, tail_gid, ntiles, n_dev, kTile);
		VGPU_HIP(hipGetLastError());
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

template <typename KeyT>
void DispatchNP(const SortedRun &run, uint64_t n, const PayloadPtrs &pay, const CompDesc &cd, const AggOutDesc &od,
                const KeyOutDesc &ko, u64 *groups_device, Scratch &scratch, hipStream_t stream,
                const uint64_t *n_dev) {
	if (pay.n == 2 && SigMatches<SigCountSumMinMaxDd>(cd)) {
		LaunchReduce<KeyT, 2, SigCountSumMinMaxDd>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
		return;
	}
	if (pay.n == 2 && SigMatches<SigCountSumMinMaxDdN>(cd)) {
		LaunchReduce<KeyT, 2, SigCountSumMinMaxDdN>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
		return;
	}
	if (pay.n == 1 && SigMatches<SigCountSumI32>(cd)) {
		LaunchReduce<KeyT, 1, SigCountSumI32>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
		return;
	}
	if (pay.n == 1 && SigMatches<SigCountSumI64>(cd)) {
		LaunchReduce<KeyT, 1, SigCountSumI64>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
		return;
	}
	if (pay.n == 1 && SigMatches<SigCountSumDd>(cd)) {
		LaunchReduce<KeyT, 1, SigCountSumDd>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
		return;
	}
	if (pay.n == 0) {
		LaunchReduce<KeyT, 0>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
	} else if (pay.n == 1) {
		LaunchReduce<KeyT, 1>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
	} else if (pay.n == 2) {
		LaunchReduce<KeyT, 2>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
	} else if (pay.n <= 4) {
		LaunchReduce<KeyT, 4>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
	} else {
		throw HipError("segmented reduce: at most 4 payload columns");
	}
}

} // namespace

void SegmentedReduce(const SortedRun &run, uint64_t n, const ReduceSpec &spec, uint64_t *groups_device,
                     Scratch &scratch, hipStream_t stream, const uint64_t *n_dev) {
	if (spec.ncomp > kMaxComponents || spec.naggs > 16) {
		throw HipError("segmented reduce: too many aggregates");
	}
	if (n == 0) {
		FillU64(groups_device, 0ull, 1, stream);
		return;
	}
	PayloadPtrs pay;
	pay.n = run.npayload;
	for (int p = 0; p < run.npayload; p++) {
		pay.ptr[p] = run.payload[p];
		pay.width[p] = run.payload_width[p];
	}
	CompDesc cd {};
	cd.ncomp = spec.ncomp;
	for (int c = 0; c < spec.ncomp; c++) {
		cd.kind[c] = spec.comp[c].kind;
		cd.payload[c] = spec.comp[c].payload;
		cd.in_type[c] = spec.comp[c].in_type;
	}
	AggOutDesc od;
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
	KeyOutDesc ko;
	ko.pack = spec.pack;
	for (int c = 0; c < spec.pack.ncols; c++) {
		ko.out[c] = spec.key_out[c];
	}
	bool count_only = run.keys != nullptr && spec.ncomp == 1 && spec.comp[0].kind == COMP_COUNT && spec.naggs > 0;
	for (int i = 0; i < spec.naggs; i++) {
		count_only = count_only && (spec.aggs[i].kind == AGG_COUNT_STAR || spec.aggs[i].kind == AGG_COUNT) &&
		             spec.aggs[i].out_type == VT_INT64;
	}
	if (count_only) {
		// Groups never outnumber rows, so n bounds the first-row buffer when no capacity is given.
		const uint64_t cap = spec.capacity == 0 ? n : std::min<uint64_t>(spec.capacity, n);
		if (run.keys64) {
			LaunchRle<u64>(run, n, cap, od, ko, groups_device, scratch, stream, n_dev);
		} else {
			LaunchRle<uint32_t>(run, n, cap, od, ko, groups_device, scratch, stream, n_dev);
		}
		return;
	}
	if (run.keys64) {
		DispatchNP<u64>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
	} else {
		DispatchNP<uint32_t>(run, n, pay, cd, od, ko, groups_device, scratch, stream, n_dev);
	}
}

} // namespace vgpu
