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
#include "segreduce_parts.h"

#include <algorithm>

namespace vgpu {
namespace {

constexpr int kRedBlock = 256;
constexpr int kRedItems = 8;
constexpr int kRedTile = kRedBlock * kRedItems;
constexpr int kRedWaves = kRedBlock / kWave;

struct PayloadPtrs {
	int n = 0;
	const void *ptr[kMaxPayload] = {};
	int width[kMaxPayload] = {};
};

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
//! tile, read in steps of 256 consecutive rows (lane l holds rows 4l .. 4l + 3 of a step: one 16-byte load
//! of 32-bit keys, so each load instruction covers a contiguous 1 KiB); no barriers.
constexpr int kHeadLaneRows = 32;
template <typename KeyT>
__global__ void __launch_bounds__(256) HeadCountKernel(const KeyT *keys, uint64_t n, uint64_t tile_rows,
                                                      uint64_t ntiles, uint32_t *counts, const uint64_t *n_dev) {
	constexpr int Q = 4;
	constexpr int kSteps = kHeadLaneRows / Q;
	if (n_dev != nullptr) { // row count written by an earlier kernel; n was its bound
		n = *n_dev;
		ntiles = (n + tile_rows - 1) / tile_rows;
	}
	const uint64_t tile = (uint64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
	if (tile >= ntiles) {
		return;
	}
	const int lane = LaneId();
	uint32_t hc = 0;
	if (keys == nullptr) {
		hc = (tile == 0 && lane == 0 && n > 0) ? 1u : 0u;
	} else {
		// tile_rows = 64 lanes x kHeadLaneRows rows
		const uint64_t tile_base = tile * tile_rows;
		const bool vec = (reinterpret_cast<uintptr_t>(keys) & 15) == 0;
		KeyT last = tile_base > 0 && tile_base <= n ? keys[tile_base - 1] : (KeyT)0;
#pragma unroll
		for (int j = 0; j < kSteps; j++) {
			const uint64_t row0 = tile_base + (uint64_t)j * (kWave * Q) + (uint64_t)lane * Q;
			KeyT k[Q];
			if (vec && row0 + Q <= n) {
				constexpr int kVecs = (int)(Q * sizeof(KeyT) / 16);
				uint4 v[kVecs];
#pragma unroll
				for (int i = 0; i < kVecs; i++) {
					v[i] = reinterpret_cast<const uint4 *>(keys + row0)[i];
				}
				__builtin_memcpy(k, v, sizeof(k));
			} else {
#pragma unroll
				for (int q = 0; q < Q; q++) {
					k[q] = row0 + q < n ? keys[row0 + q] : (KeyT)0;
				}
			}
			const KeyT up = __shfl_up(k[Q - 1], 1);
			KeyT before = lane == 0 ? last : up;
#pragma unroll
			for (int q = 0; q < Q; q++) {
				hc += (row0 + q < n && (row0 + q == 0 || k[q] != before)) ? 1u : 0u;
				before = k[q];
			}
			last = __shfl(k[Q - 1], kWave - 1);
		}
	}
	for (int off = kWave / 2; off > 0; off >>= 1) {
		hc += __shfl_xor(hc, off);
	}
	if (lane == 0) {
		counts[tile] = hc;
	}
}

//! Count-only reduce (every aggregate is a COUNT): a run-length encoding of the sorted keys. One wave per
//! 2,048-row tile (the head-count tiles, so the tile's first group id comes from the head-count scan and no
//! tile takes a ticket or waits on another). The tile is read in 8 steps of 256 consecutive rows, lane l
//! holding rows 4l .. 4l + 3 of the step (one 16-byte load of 32-bit keys); a row's predecessor is the
//! previous lane's last key (the previous step's last key for lane 0). A head's group id is the tile's first
//! id + the heads before it: four ballots per step give the heads of the lower lanes, the lane's own bits
//! those before it in the lane. Each group's key and first row are written in row order, so the stores of
//! one step land in a few cache lines; RleCountKernel turns first rows into counts.
template <typename KeyT>
__global__ void __launch_bounds__(256) RleHeadsKernel(const KeyT *keys, uint64_t n, KeyOutDesc ko, u64 *starts,
                                                     const u64 *tile_prefix, const u64 *tile_carry,
                                                     uint64_t ntiles, const uint64_t *n_dev) {
	constexpr int Q = 4;
	constexpr int kSteps = kHeadLaneRows / Q;
	constexpr uint64_t kTileRows = (uint64_t)kWave * kHeadLaneRows;
	if (n_dev != nullptr) {
		n = *n_dev;
		ntiles = (n + kTileRows - 1) / kTileRows;
	}
	const uint64_t tile = (uint64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
	if (tile >= ntiles) {
		return;
	}
	const int lane = LaneId();
	const uint64_t tile_base = tile * kTileRows;
	const bool vec = (reinterpret_cast<uintptr_t>(keys) & 15) == 0;
	u64 gid = tile_prefix[tile] + tile_carry[tile / kScanTiles];
	KeyT last = tile_base > 0 && tile_base <= n ? keys[tile_base - 1] : (KeyT)0; // the row before the step
#pragma unroll
	for (int j = 0; j < kSteps; j++) {
		const uint64_t row0 = tile_base + (uint64_t)j * (kWave * Q) + (uint64_t)lane * Q;
		KeyT k[Q];
		if (vec && row0 + Q <= n) {
			// 4 consecutive keys: 16 B (u32) or 32 B (u64), 16-byte aligned (row0 is a multiple of 4)
			constexpr int kVecs = (int)(Q * sizeof(KeyT) / 16);
			uint4 v[kVecs];
#pragma unroll
			for (int i = 0; i < kVecs; i++) {
				v[i] = reinterpret_cast<const uint4 *>(keys + row0)[i];
			}
			__builtin_memcpy(k, v, sizeof(k));
		} else {
#pragma unroll
			for (int q = 0; q < Q; q++) {
				k[q] = row0 + q < n ? keys[row0 + q] : (KeyT)0;
			}
		}
		const KeyT up = __shfl_up(k[Q - 1], 1);
		KeyT before = lane == 0 ? last : up;
		uint32_t h = 0; // bit q: row0 + q starts a group
#pragma unroll
		for (int q = 0; q < Q; q++) {
			const bool head = row0 + q < n && (row0 + q == 0 || k[q] != before);
			before = k[q];
			h |= head ? (1u << q) : 0u;
		}
		uint32_t below = 0, total = 0;
#pragma unroll
		for (int q = 0; q < Q; q++) {
			const u64 b = __ballot((h >> q) & 1u);
			below += __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
			total += (uint32_t)Popc64(b);
		}
		u64 g = gid + below;
#pragma unroll
		for (int q = 0; q < Q; q++) {
			if ((h >> q) & 1u) {
				WriteKeys(ko, (u64)k[q], g);
				starts[g] = row0 + q;
				g++;
			}
		}
		gid += total;
		last = __shfl(k[Q - 1], kWave - 1);
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
	CompDesc cd;
	AggOutDesc od;
	KeyOutDesc ko;
	BuildReduceDescs(spec, cd, od, ko);
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
