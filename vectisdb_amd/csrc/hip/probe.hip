// probe.hip -- join-aggregate probe of many row ranges against a resident build-side index (gfx950).
//
// The index is the build side reduced to its distinct keys in ascending order, each with its row count
// and the exact int128 sum of its payload (the GROUP BY operator builds it once and it stays in HBM).
// A probe row with key x contributes count[x] rows and sum[x] to COUNT(*) and SUM(payload) of the join,
// so SUM/COUNT of an INNER equi-join over a probe range need no sort of the probe side: one pass over the
// probe keys, a lower_bound per row, and a per-range reduction.
//
// The ranges are laid out like the fused kernel's row ranges: range g covers launch rows
// [seg[g], seg[g] + seg_len[g]), read from table rows seg_row[g] + (v - seg[g]); other launch rows are
// padding. A tile of kItems x 256 launch rows usually lies inside one range: its rows are summed in
// registers, and a block carries those sums over its consecutive tiles of the same range, adding them with
// three 64-bit atomics when the range changes; a tile that straddles ranges adds row by row.
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <type_traits>

namespace vgpu {
namespace {

constexpr int kProbeBlock = 256;
constexpr int kProbeItems = 8;
constexpr uint64_t kProbeTile = (uint64_t)kProbeBlock * kProbeItems;
constexpr uint64_t kProbeBlocks = 2048;

//! Index of the range that holds launch row v: the last g with seg[g] <= v (empty ranges skipped).
__device__ __forceinline__ uint32_t RangeOf(const u64 *seg, uint32_t lo, uint32_t hi, u64 v) {
	// first g in [lo, hi) with seg[g + 1] > v
	while (lo < hi) {
		const uint32_t mid = (lo + hi) >> 1;
		if (seg[mid + 1] <= v) {
			lo = mid + 1;
		} else {
			hi = mid;
		}
	}
	return lo;
}

__device__ __forceinline__ void Add128(u64 &lo, u64 &hi, u64 xlo, u64 xhi) {
	const u64 s = lo + xlo;
	hi += xhi + (s < lo ? 1ull : 0ull);
	lo = s;
}

__device__ __forceinline__ void AtomicAdd128(u64 *p, u64 lo, u64 hi) {
	const u64 old = atomicAdd((unsigned long long *)p, (unsigned long long)lo);
	const u64 carry = (old + lo < old) ? 1ull : 0ull;
	atomicAdd((unsigned long long *)(p + 1), (unsigned long long)(hi + carry));
}

//! Dense entry: row count in the top 23 bits, the payload sum as a 41-bit two's complement below.
constexpr int kDenseSumBits = 41;

constexpr uint32_t kNoRange = 0xffffffffu;

//! Adds the block's partial sums for range `g` to acc (block-wide: every thread calls it with the same g;
//! g = kNoRange adds nothing).
__device__ __forceinline__ void FlushRange(uint32_t g, u64 c, u64 slo, u64 shi, u64 (*s_red)[kProbeBlock / kWave],
                                           u64 *acc) {
	if (g == kNoRange) {
		return;
	}
	for (int d = kWave / 2; d > 0; d >>= 1) {
		const u64 oc = __shfl_xor(c, d);
		const u64 olo = __shfl_xor(slo, d);
		const u64 ohi = __shfl_xor(shi, d);
		c += oc;
		Add128(slo, shi, olo, ohi);
	}
	const int tid = threadIdx.x;
	const int wave = tid / kWave;
	if ((tid & (kWave - 1)) == 0) {
		s_red[0][wave] = c;
		s_red[1][wave] = slo;
		s_red[2][wave] = shi;
	}
	__syncthreads();
	if (tid == 0) {
		u64 tc = 0, tlo = 0, thi = 0;
		for (int w = 0; w < kProbeBlock / kWave; w++) {
			tc += s_red[0][w];
			Add128(tlo, thi, s_red[1][w], s_red[2][w]);
		}
		if (tc != 0) {
			u64 *a = acc + 3 * (u64)g;
			atomicAdd((unsigned long long *)a, (unsigned long long)tc);
			AtomicAdd128(a + 1, tlo, thi);
		}
	}
	__syncthreads();
}

//! Each block walks `tiles_per_block` consecutive tiles. While consecutive tiles lie inside one range the
//! block keeps summing in registers and adds to that range's accumulator once, when the range changes or
//! the block ends (a long range costs one set of atomics per block, not per tile); a tile that straddles
//! ranges adds row by row.
template <typename K>
__global__ void __launch_bounds__(kProbeBlock)
    RangeProbeAggKernel(const K *fk, const u64 *seg, const u64 *seg_row, const u64 *seg_len, uint32_t nseg,
                        u64 nvirt, u64 tiles_per_block, const K *dkeys, const u64 *dcount, const u64 *dsum, u64 nd,
                        const u64 *dense, u64 dense_min, u64 dense_span, u64 *acc) {
	__shared__ uint32_t s_g[2];
	__shared__ u64 s_red[3][kProbeBlock / kWave];
	const int tid = threadIdx.x;
	const u64 ntiles = (nvirt + kProbeTile - 1) / kProbeTile;
	const u64 t_begin = (u64)blockIdx.x * tiles_per_block;
	const u64 t_end = t_begin + tiles_per_block < ntiles ? t_begin + tiles_per_block : ntiles;
	u64 c = 0, slo = 0, shi = 0; // this thread's sums for range `cur`
	uint32_t cur = kNoRange;
	for (u64 t = t_begin; t < t_end; t++) {
		const u64 tile_base = t * kProbeTile;
		const u64 tile_last = (tile_base + kProbeTile - 1 < nvirt ? tile_base + kProbeTile : nvirt) - 1;
		__syncthreads(); // every thread has read the previous tile's s_g
		if (tid == 0) {
			s_g[0] = RangeOf(seg, 0, nseg, tile_base);
			s_g[1] = RangeOf(seg, 0, nseg, tile_last);
		}
		__syncthreads();
		const uint32_t g0 = s_g[0], g1 = s_g[1];
		if (g0 == g1 && g0 != cur) {
			FlushRange(cur, c, slo, shi, s_red, acc);
			c = 0;
			slo = 0;
			shi = 0;
			cur = g0;
		}
		K x[kProbeItems];
		u64 pos[kProbeItems];
		uint32_t gi[kProbeItems];
		bool live[kProbeItems];
#pragma unroll
		for (int k = 0; k < kProbeItems; k++) {
			const u64 v = tile_base + (u64)k * kProbeBlock + tid;
			// v <= tile_last lies in ranges g0..g1, so the search never passes g1
			const uint32_t g = (g0 == g1 || v > tile_last) ? g0 : RangeOf(seg, g0, g1 + 1, v);
			const u64 off = v - seg[g];
			live[k] = v <= tile_last && off < seg_len[g];
			gi[k] = g;
			x[k] = live[k] ? fk[seg_row[g] + off] : (K)0;
			pos[k] = 0;
		}
		if (dense != nullptr) {
			// Direct index: one load per probe row (the table is small enough to stay in the Infinity Cache).
			u64 e[kProbeItems];
#pragma unroll
			for (int k = 0; k < kProbeItems; k++) {
				const u64 off = (u64)(long long)x[k] - dense_min; // wraps past dense_span when out of range
				e[k] = (live[k] && off < dense_span) ? dense[off] : 0ull;
			}
#pragma unroll
			for (int k = 0; k < kProbeItems; k++) {
				const u64 n = e[k] >> kDenseSumBits;
				if (n == 0) {
					continue;
				}
				const long long sv = (long long)(e[k] << (64 - kDenseSumBits)) >> (64 - kDenseSumBits);
				const u64 ylo = (u64)sv, yhi = sv < 0 ? ~0ull : 0ull;
				if (g0 == g1) {
					c += n;
					Add128(slo, shi, ylo, yhi);
				} else {
					u64 *a = acc + 3 * (u64)gi[k];
					atomicAdd((unsigned long long *)a, (unsigned long long)n);
					AtomicAdd128(a + 1, ylo, yhi);
				}
			}
		} else {
			// Branch-free lower_bound of all items at once: every item halves the same span per step, so the
			// loads of one step are independent and in flight together.
			if (nd > 0) {
				u64 len = nd;
				while (len > 1) {
					const u64 half = len >> 1;
#pragma unroll
					for (int k = 0; k < kProbeItems; k++) {
						pos[k] = dkeys[pos[k] + half] < x[k] ? pos[k] + half : pos[k];
					}
					len -= half;
				}
#pragma unroll
				for (int k = 0; k < kProbeItems; k++) {
					pos[k] += dkeys[pos[k]] < x[k] ? 1ull : 0ull;
				}
			}
#pragma unroll
			for (int k = 0; k < kProbeItems; k++) {
				const bool hit = live[k] && pos[k] < nd && dkeys[pos[k]] == x[k];
				if (!hit) {
					continue;
				}
				const u64 n = dcount[pos[k]];
				const u64 ylo = dsum[2 * pos[k]], yhi = dsum[2 * pos[k] + 1];
				if (g0 == g1) {
					c += n;
					Add128(slo, shi, ylo, yhi);
				} else {
					u64 *a = acc + 3 * (u64)gi[k];
					atomicAdd((unsigned long long *)a, (unsigned long long)n);
					AtomicAdd128(a + 1, ylo, yhi);
				}
			}
		}
	}
	FlushRange(cur, c, slo, shi, s_red, acc);
}

struct DenseArgs {
	const u64 *table;
	u64 min;
	u64 span;
};

template <typename K>
void Launch(const void *fk, const uint64_t *seg, const uint64_t *seg_row, const uint64_t *seg_len, uint32_t nseg,
            uint64_t nvirt, const void *dkeys, const uint64_t *dcount, const uint64_t *dsum, uint64_t nd,
            const DenseArgs &dense, uint64_t *acc, hipStream_t stream) {
	const uint64_t tiles = (nvirt + kProbeTile - 1) / kProbeTile;
	// Up to kProbeBlocks blocks (8 waves per SIMD on every CU), each over a contiguous run of tiles.
	const uint64_t blocks = std::min<uint64_t>(tiles, kProbeBlocks);
	const uint64_t per_block = (tiles + blocks - 1) / blocks;
	hipLaunchKernelGGL((RangeProbeAggKernel<K>), dim3((unsigned)blocks), dim3(kProbeBlock), 0, stream,
	                   (const K *)fk, seg, seg_row, seg_len, nseg, nvirt, (u64)per_block, (const K *)dkeys, dcount,
	                   dsum, nd, dense.table, dense.min, dense.span, acc);
	VGPU_HIP(hipGetLastError());
}

template <typename K>
__global__ void DenseIndexKernel(const K *keys, const u64 *count, const u64 *sum, u64 nd, u64 min, u64 *dense,
                                 unsigned int *overflow) {
	const u64 stride = (u64)gridDim.x * blockDim.x;
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += stride) {
		const u64 c = count[i];
		const u64 lo = sum[2 * i], hi = sum[2 * i + 1];
		const long long s = (long long)lo;
		const bool fits = c < (1ull << (64 - kDenseSumBits)) && hi == (s < 0 ? ~0ull : 0ull) &&
		                  s >= -(1ll << (kDenseSumBits - 1)) && s < (1ll << (kDenseSumBits - 1));
		if (!fits) {
			atomicOr(overflow, 1u);
		}
		const u64 mask = (1ull << kDenseSumBits) - 1;
		dense[(u64)(long long)keys[i] - min] = (c << kDenseSumBits) | (lo & mask);
	}
}

// ---------------------------------------------------------------- partitioned probe (one large probe side)
//
// A probe side much larger than the build side reads the direct-address table at random: every lookup
// fetches a whole cache line over the fabric into an L2 that holds 1/32 of a 128 MB table, so the probe is
// bound by line traffic, not by HBM. Partitioned, the probe keys are grouped by key range into 256 buckets
// (each bucket's slice of the table is 1/256 of it), and the probe reads one slice at a time. Steps:
//   1. PartBatchKernel: block g takes batch g (32768 consecutive probe rows), drops the keys outside the
//      table, groups the rest by bucket in LDS (LDS atomics give each key its rank in its bucket, a scan of
//      the 256 counts the buckets' starts) and writes them, as table offsets (uint32), over the batch's own
//      positions of `part`: one contiguous write per batch, and tab[g] = the 257 bucket starts. Order inside
//      a bucket does not matter for SUM and COUNT, and no global histogram pass is needed;
//   2. PartProbeKernel<LDS>: block (b, split) walks bucket b's run in every batch of its split (a scan of the
//      run lengths in LDS maps consecutive threads to consecutive keys of a run), looking each key up in
//      the bucket's table slice -- copied to LDS when the table has a 16-bit form and the slice fits, read
//      from L2 otherwise -- and sums count and payload sum per block; PartFinalKernel adds the partials.
constexpr int kPartBuckets = 256;
constexpr int kPartBlock = 1024;      // 16 waves
constexpr int kPartItems = 32;        // probe: keys per thread per loop step (loads in flight per lane)
constexpr int kPartItemsY = 16;       // the same with a payload (32 would spill registers)
constexpr int kScatterItems = 32;     // keys per thread per batch (keys only: 32768-row batches)
constexpr int kScatterItemsY = 16;    // with a probe payload (16384-row batches)
constexpr unsigned kRunChunk = 2048;  // runs (batches) per LDS scan of run lengths

//! Raw integer of `width` bytes at row j, sign-extended (width 1 is BOOLEAN: zero-extended).
__device__ __forceinline__ long long LoadIntAt(const void *p, int width, u64 j) {
	switch (width) {
	case 1:
		return ((const uint8_t *)p)[j];
	case 2:
		return ((const int16_t *)p)[j];
	case 4:
		return ((const int32_t *)p)[j];
	default:
		return ((const long long *)p)[j];
	}
}

//! Batch g = rows [g * ITEMS * 1024, ...): kept keys grouped by bucket, written over the batch's own positions
//! of `part` (table offsets) and, with a payload type Y (int32 for payloads of up to 4 bytes, int64 for 8;
//! void: none), of `part_y` (the probe payload, same order).
template <typename K, int ITEMS, typename Y>
__global__ void __launch_bounds__(kPartBlock) PartBatchKernel(const K *keys, u64 n, u64 dmin, u64 span, int shift,
                                                             const void *y, int ywidth, uint32_t *part, void *part_y,
                                                             uint32_t *tab) {
	constexpr bool HASY = !std::is_void<Y>::value;
	constexpr int kYBytes = HASY ? (int)sizeof(typename std::conditional<HASY, Y, int>::type) : 4;
	constexpr u64 kBatch = (u64)kPartBlock * ITEMS;
	__shared__ uint32_t s_cnt[kPartBuckets];   // batch count per bucket
	__shared__ uint32_t s_start[kPartBuckets]; // batch-local start per bucket
	__shared__ uint32_t s_scan[kPartBlock / kWave + 1];
	__shared__ u64 s_stage[kBatch * (kYBytes > 4 ? kYBytes : 4) / 8]; // offsets, then payloads
	uint32_t *s_off = reinterpret_cast<uint32_t *>(s_stage);
	const int tid = threadIdx.x;
	const u64 base = (u64)blockIdx.x * kBatch;
	uint32_t off[ITEMS], rank[ITEMS];
	bool valid[ITEMS];
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		const u64 r = base + (u64)j * kPartBlock + tid;
		const K x = r < n ? keys[r] : (K)0;
		const u64 o = (u64)(long long)x - dmin; // wraps past span when below the table
		valid[j] = r < n && o < span;
		off[j] = (uint32_t)o;
	}
	if (tid < kPartBuckets) {
		s_cnt[tid] = 0;
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		rank[j] = valid[j] ? atomicAdd(&s_cnt[off[j] >> shift], 1u) : 0u;
	}
	__syncthreads();
	uint32_t total = 0;
	const uint32_t excl =
	    BlockExclusiveScan<uint32_t, kPartBlock>(tid < kPartBuckets ? s_cnt[tid] : 0u, s_scan, &total);
	uint32_t *row = tab + (u64)blockIdx.x * (kPartBuckets + 1);
	if (tid < kPartBuckets) {
		s_start[tid] = excl;
		row[tid] = excl;
	}
	if (tid == 0) {
		row[kPartBuckets] = total;
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		if (valid[j]) {
			s_off[s_start[off[j] >> shift] + rank[j]] = off[j];
		}
	}
	__syncthreads();
	for (uint32_t i = tid; i < total; i += kPartBlock) {
		part[base + i] = s_off[i];
	}
	if constexpr (HASY) {
		using YT = typename std::conditional<HASY, Y, int>::type;
		YT *s_y = reinterpret_cast<YT *>(s_stage);
		YT yv[ITEMS];
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			yv[j] = valid[j] ? (YT)LoadIntAt(y, ywidth, base + (u64)j * kPartBlock + tid) : (YT)0;
		}
		__syncthreads(); // every offset has left the stage
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			if (valid[j]) {
				s_y[s_start[off[j] >> shift] + rank[j]] = yv[j];
			}
		}
		__syncthreads();
		for (uint32_t i = tid; i < total; i += kPartBlock) {
			static_cast<YT *>(part_y)[base + i] = s_y[i];
		}
	}
}

//! Adds one dense entry (count in the top 23 bits, 41-bit two's complement sum) to (c, slo:shi).
__device__ __forceinline__ void AddDense(u64 e, u64 &c, u64 &slo, u64 &shi) {
	const long long sv = (long long)(e << (64 - kDenseSumBits)) >> (64 - kDenseSumBits);
	c += e >> kDenseSumBits;
	Add128(slo, shi, (u64)sv, sv < 0 ? ~0ull : 0ull);
}

// 16-bit form of the direct-address table: count in the top 4 bits, the payload sum as a 12-bit two's
// complement below (build keys that are nearly unique with small payloads). With 256 buckets over a span
// of at most 2^24 a bucket's slice is at most 64K entries = 128 KB, which fits in LDS (160 KB per CU on
// gfx950), so the probe reads the table from LDS instead of L2.
constexpr int kNarrowSumBits = 12;
constexpr u64 kSliceEntries = 65536;
constexpr unsigned kSliceSplits = 4; // blocks per bucket

__global__ void __launch_bounds__(256) NarrowDenseKernel(const u64 *dense, u64 span, uint16_t *out,
                                                        unsigned int *nofit) {
	const u64 stride = (u64)gridDim.x * blockDim.x;
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < span; i += stride) {
		const u64 e = dense[i];
		const u64 c = e >> kDenseSumBits;
		const long long sv = (long long)(e << (64 - kDenseSumBits)) >> (64 - kDenseSumBits);
		const bool fits = c < 16 && sv >= -(1ll << (kNarrowSumBits - 1)) && sv < (1ll << (kNarrowSumBits - 1));
		out[i] = (uint16_t)((c << kNarrowSumBits) | ((u64)sv & ((1ull << kNarrowSumBits) - 1)));
		if (__ballot(!fits) != 0 && LaneId() == 0 &&
		    __hip_atomic_load(nofit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
			atomicOr(nofit, 1u);
		}
	}
}

//! y x c as int128 for a signed y and a count c < 2^63, added to (lo, hi).
__device__ __forceinline__ void AddMul128(u64 &lo, u64 &hi, long long y, u64 c) {
	const u64 plo = (u64)y * c;
	const u64 phi = __umul64hi((u64)y, c) - (y < 0 ? c : 0ull); // (y + 2^64) * c overcounts 2^64 * c
	Add128(lo, hi, plo, phi);
}

//! Block i probes split i % kSliceSplits (a range of batches) of bucket i / kSliceSplits. LDS: the bucket's
//! 16-bit table slice is copied to LDS once and every key is looked up there; otherwise the 64-bit table is
//! read through L2. Sums per block: matched build rows, the build payload sum, and (HASY) the probe payload
//! times its key's build row count.
template <bool LDS, typename Y>
__global__ void __launch_bounds__(kPartBlock) PartProbeKernel(const uint32_t *part, const void *part_y,
                                                             const uint32_t *tab, u64 nbatches, u64 batch,
                                                             const uint16_t *dense16, const u64 *dense, u64 span,
                                                             int shift, u64 *partials) {
	constexpr bool HASY = !std::is_void<Y>::value;
	using YT = typename std::conditional<HASY, Y, int>::type;
	constexpr int kItems = HASY ? kPartItemsY : kPartItems;
	__shared__ __align__(16) uint16_t s_slice[LDS ? kSliceEntries : 1];
	__shared__ uint32_t s_pre[kRunChunk + 1]; // exclusive prefix of run lengths
	__shared__ uint32_t s_first[kRunChunk];   // each run's start inside its batch
	__shared__ uint32_t s_scan[kPartBlock / kWave + 1];
	__shared__ u64 s_red[5][kPartBlock / kWave];
	const int tid = threadIdx.x;
	const unsigned b = blockIdx.x / kSliceSplits;
	const unsigned split = blockIdx.x % kSliceSplits;
	const u64 e0 = (u64)b << shift;
	if constexpr (LDS) {
		const u64 slice = 1ull << shift;
		const u64 ne = e0 < span ? (span - e0 < slice ? span - e0 : slice) : 0;
		// 16-byte pieces when the slice starts 16-byte aligned (every slice of an aligned table does), then
		// the tail entry by entry
		const u64 nv = (reinterpret_cast<uintptr_t>(dense16 + e0) & 15) == 0 ? ne / 8 : 0;
		const uint4 *src = reinterpret_cast<const uint4 *>(dense16 + e0);
		uint4 *dst = reinterpret_cast<uint4 *>(s_slice);
		for (u64 i = tid; i < nv; i += kPartBlock) {
			dst[i] = src[i];
		}
		for (u64 i = nv * 8 + tid; i < ne; i += kPartBlock) {
			s_slice[i] = dense16[e0 + i];
		}
	}
	const u64 per = (nbatches + kSliceSplits - 1) / kSliceSplits;
	const u64 g_begin = (u64)split * per < nbatches ? (u64)split * per : nbatches;
	const u64 g_end = g_begin + per < nbatches ? g_begin + per : nbatches;
	u64 c = 0, slo = 0, shi = 0, ylo = 0, yhi = 0;
	for (u64 g0 = g_begin; g0 < g_end; g0 += kRunChunk) {
		const unsigned nruns = (unsigned)(g_end - g0 < kRunChunk ? g_end - g0 : kRunChunk);
		// run lengths of this chunk of batches: two runs per thread, scanned
		uint32_t len[2], first[2];
#pragma unroll
		for (int k = 0; k < 2; k++) {
			const unsigned i = tid * 2 + k;
			const uint32_t *row = tab + (g0 + i) * (kPartBuckets + 1);
			first[k] = i < nruns ? row[b] : 0u;
			len[k] = i < nruns ? row[b + 1] - first[k] : 0u;
		}
		uint32_t total = 0;
		const uint32_t excl = BlockExclusiveScan<uint32_t, kPartBlock>(len[0] + len[1], s_scan, &total);
		s_pre[tid * 2] = excl;
		s_pre[tid * 2 + 1] = excl + len[0];
		s_first[tid * 2] = first[0];
		s_first[tid * 2 + 1] = first[1];
		if (tid == 0) {
			s_pre[kRunChunk] = total;
		}
		__syncthreads(); // also orders the slice copy before the first lookup
		// Wave w takes 64 x kItems consecutive keys of the chunk's concatenated runs per step, lane l the
		// keys l, l + 64, ... of them: loads are coalesced within a run. Each lane finds its first key's run
		// by binary search (the lanes of a wave follow nearly the same path) and steps forward from there.
		const uint32_t wave = (uint32_t)tid / kWave, lane = (uint32_t)tid % kWave;
		for (uint32_t vb = wave * kWave * kItems; vb < total; vb += kPartBlock * kItems) {
			const uint32_t v0 = vb + lane;
			unsigned lo = 0, hi = nruns; // invariant: s_pre[lo] <= min(v0, total - 1) < s_pre[hi]
			const uint32_t vs = v0 < total ? v0 : total - 1;
			while (hi - lo > 1) {
				const unsigned mid = (lo + hi) >> 1;
				if (s_pre[mid] <= vs) {
					lo = mid;
				} else {
					hi = mid;
				}
			}
			uint32_t key[kItems];
			long long yv[kItems];
#pragma unroll
			for (int k = 0; k < kItems; k++) {
				const uint32_t v = v0 + k * kWave;
				key[k] = (uint32_t)e0; // any offset of this bucket
				yv[k] = 0;
				if (v < total) {
					while (s_pre[lo + 1] <= v) { // v < total = s_pre[nruns], so lo stays below nruns
						lo++;
					}
					const u64 at = (g0 + lo) * batch + s_first[lo] + (v - s_pre[lo]);
					key[k] = part[at];
					if constexpr (HASY) {
						yv[k] = (long long)static_cast<const YT *>(part_y)[at];
					}
				}
			}
#pragma unroll
			for (int k = 0; k < kItems; k++) {
				if (v0 + k * kWave < total) {
					u64 ec;
					long long es;
					if constexpr (LDS) {
						const uint32_t e = s_slice[key[k] - (uint32_t)e0];
						ec = e >> kNarrowSumBits;
						es = (long long)((int)(e << (32 - kNarrowSumBits)) >> (32 - kNarrowSumBits));
					} else {
						const u64 e = dense[key[k]];
						ec = e >> kDenseSumBits;
						es = (long long)(e << (64 - kDenseSumBits)) >> (64 - kDenseSumBits);
					}
					c += ec;
					Add128(slo, shi, (u64)es, es < 0 ? ~0ull : 0ull);
					if constexpr (HASY) {
						AddMul128(ylo, yhi, yv[k], ec);
					}
				}
			}
		}
		__syncthreads(); // s_pre / s_first are rewritten by the next chunk
	}
	for (int d = kWave / 2; d > 0; d >>= 1) {
		const u64 oc = __shfl_xor(c, d);
		const u64 olo = __shfl_xor(slo, d);
		const u64 ohi = __shfl_xor(shi, d);
		const u64 oylo = __shfl_xor(ylo, d);
		const u64 oyhi = __shfl_xor(yhi, d);
		c += oc;
		Add128(slo, shi, olo, ohi);
		Add128(ylo, yhi, oylo, oyhi);
	}
	const int wave = tid / kWave;
	if ((tid & (kWave - 1)) == 0) {
		s_red[0][wave] = c;
		s_red[1][wave] = slo;
		s_red[2][wave] = shi;
		s_red[3][wave] = ylo;
		s_red[4][wave] = yhi;
	}
	__syncthreads();
	if (tid == 0) {
		u64 tc = 0, tlo = 0, thi = 0, tylo = 0, tyhi = 0;
		for (int w = 0; w < kPartBlock / kWave; w++) {
			tc += s_red[0][w];
			Add128(tlo, thi, s_red[1][w], s_red[2][w]);
			Add128(tylo, tyhi, s_red[3][w], s_red[4][w]);
		}
		u64 *out = partials + 5 * (u64)blockIdx.x;
		out[0] = tc;
		out[1] = tlo;
		out[2] = thi;
		out[3] = tylo;
		out[4] = tyhi;
	}
}

//! One wave adds the block partials (integers: any order gives the same result).
__global__ void __launch_bounds__(kWave) PartFinalKernel(const u64 *partials, unsigned nparts, u64 *acc) {
	u64 c = 0, slo = 0, shi = 0, ylo = 0, yhi = 0;
	for (unsigned p = LaneId(); p < nparts; p += kWave) {
		const u64 *in = partials + 5 * (u64)p;
		c += in[0];
		Add128(slo, shi, in[1], in[2]);
		Add128(ylo, yhi, in[3], in[4]);
	}
	for (int d = kWave / 2; d > 0; d >>= 1) {
		const u64 oc = __shfl_xor(c, d);
		const u64 olo = __shfl_xor(slo, d);
		const u64 ohi = __shfl_xor(shi, d);
		const u64 oylo = __shfl_xor(ylo, d);
		const u64 oyhi = __shfl_xor(yhi, d);
		c += oc;
		Add128(slo, shi, olo, ohi);
		Add128(ylo, yhi, oylo, oyhi);
	}
	if (LaneId() == 0) {
		acc[0] = c;
		acc[1] = slo;
		acc[2] = shi;
		acc[3] = ylo;
		acc[4] = yhi;
	}
}

//! The batch scatter of one probe side (key type K, payload Y or void).
template <typename K, typename Y>
void LaunchScatterY(const ProbePartition &p, const K *keys, const void *y, hipStream_t stream) {
	// with a payload (4 or 8 bytes) a thread holds 16 rows: 32 rows of offsets, ranks and payloads cost more in
	// registers than the longer bucket runs save
	constexpr int kItems = std::is_void<Y>::value ? kScatterItems : kScatterItemsY;
	const double ybytes = std::is_void<Y>::value ? 0.0 : (double)p.payload_width + 2.0 * sizeof(typename std::conditional<
	                                                                                std::is_void<Y>::value, int, Y>::type);
	ProfileScope prof("partition_scatter", stream, (double)p.n * (sizeof(K) + 4) + (double)p.n * ybytes / 2);
	hipLaunchKernelGGL((PartBatchKernel<K, kItems, Y>), dim3((unsigned)p.nbatches), dim3(kPartBlock), 0, stream,
	                   keys, (u64)p.n, (u64)p.dense_min, (u64)p.span, p.shift, y, p.payload_width, p.part, p.part_y,
	                   p.tab);
	VGPU_HIP(hipGetLastError());
}

template <typename K>
void LaunchScatter(const ProbePartition &p, const void *keys_v, const void *y, hipStream_t stream) {
	const K *keys = static_cast<const K *>(keys_v);
	if (y == nullptr) {
		LaunchScatterY<K, void>(p, keys, y, stream);
	} else if (p.payload_width <= 4) {
		LaunchScatterY<K, int32_t>(p, keys, y, stream);
	} else {
		LaunchScatterY<K, long long>(p, keys, y, stream);
	}
}

//! The per-bucket probe of a partitioned probe side (payload Y or void) and the final fold into acc.
template <typename Y>
void LaunchBucketProbeY(const ProbePartition &p, const uint64_t *dense, const uint16_t *dense16, uint64_t *acc,
                        hipStream_t stream) {
	const unsigned nparts = kPartBuckets * kSliceSplits;
	const double ybytes = std::is_void<Y>::value ? 0.0 : (double)p.payload_width + 2.0 * sizeof(typename std::conditional<
	                                                                                std::is_void<Y>::value, int, Y>::type);
	{
		// the two forms are profiled under their own names: "_lds" looks keys up in the bucket's 16-bit table
		// slice copied to LDS, "_l2" in the 64-bit table through the caches
		const bool lds = dense16 != nullptr && (1ull << p.shift) <= kSliceEntries;
		ProfileScope prof(lds ? "partition_probe_lds" : "partition_probe_l2", stream,
		                  (double)p.n * 4 + (double)p.n * ybytes / 2);
		if (lds) {
			hipLaunchKernelGGL((PartProbeKernel<true, Y>), dim3(nparts), dim3(kPartBlock), 0, stream, p.part,
			                   p.part_y, p.tab, (u64)p.nbatches, (u64)p.batch, dense16, dense, (u64)p.span, p.shift,
			                   p.partials);
		} else {
			hipLaunchKernelGGL((PartProbeKernel<false, Y>), dim3(nparts), dim3(kPartBlock), 0, stream, p.part,
			                   p.part_y, p.tab, (u64)p.nbatches, (u64)p.batch, dense16, dense, (u64)p.span, p.shift,
			                   p.partials);
		}
		VGPU_HIP(hipGetLastError());
	}
	hipLaunchKernelGGL(PartFinalKernel, dim3(1), dim3(kWave), 0, stream, p.partials, nparts, acc);
	VGPU_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- table straight from the sorted build side

constexpr int kRunItems = 8;        // sorted rows per thread
constexpr uint64_t kMaxRunRows = 64; // a longer run of one key is left to the GROUP BY path

//! Sorted build side (packed keys ascending, payload permuted with them) -> direct-address table entries.
//! The thread holding a run's first row sums the run (reading past its own rows when the run continues)
//! and stores the entry; every other entry stays 0. flags bit 0: an entry does not pack; bit 1: a run is
//! longer than kMaxRunRows (the caller then takes the GROUP BY path).
template <typename KeyT>
__global__ void __launch_bounds__(256) DenseFromRunKernel(const KeyT *keys, const void *payload, int pwidth,
                                                         u64 n, u64 pack_min, u64 flip, u64 dmin, u64 span,
                                                         u64 *dense, unsigned int *flags) {
	const u64 row0 = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * kRunItems;
	if (row0 >= n) {
		return;
	}
	// This thread's rows in registers (vector loads when all of them are there), and the key before them.
	KeyT kr[kRunItems];
	long long pr[kRunItems];
	const bool full = row0 + kRunItems <= n;
	if (full && (reinterpret_cast<uintptr_t>(keys + row0) & 15) == 0) {
		uint4 v[kRunItems * sizeof(KeyT) / 16];
#pragma unroll
		for (int q = 0; q < (int)(kRunItems * sizeof(KeyT) / 16); q++) {
			v[q] = reinterpret_cast<const uint4 *>(keys + row0)[q];
		}
		__builtin_memcpy(kr, v, sizeof(kr));
	} else {
#pragma unroll
		for (int i = 0; i < kRunItems; i++) {
			kr[i] = row0 + i < n ? keys[row0 + i] : (KeyT)0;
		}
	}
#pragma unroll
	for (int i = 0; i < kRunItems; i++) {
		pr[i] = (payload != nullptr && row0 + i < n) ? LoadIntAt(payload, pwidth, row0 + i) : 0ll;
	}
	const KeyT before = row0 > 0 ? keys[row0 - 1] : (KeyT)0;
	unsigned int bad = 0;
#pragma unroll
	for (int i = 0; i < kRunItems; i++) {
		const u64 r = row0 + i;
		const KeyT k = kr[i];
		const bool head = r < n && (r == 0 || (i == 0 ? before : kr[i > 0 ? i - 1 : 0]) != k);
		if (!head) {
			continue;
		}
		u64 cnt = 0, slo = 0, shi = 0;
		bool open = true; // the run reaches the end of this thread's rows
#pragma unroll
		for (int j = 0; j < kRunItems; j++) {
			if (j >= i && open) {
				if (row0 + j < n && kr[j] == k) {
					cnt++;
					Add128(slo, shi, (u64)pr[j], pr[j] < 0 ? ~0ull : 0ull);
				} else {
					open = false;
				}
			}
		}
		if (open) {
			// the run continues into the next thread's rows
			u64 j = row0 + kRunItems;
			for (; j < n && cnt < kMaxRunRows && keys[j] == k; j++) {
				if (payload != nullptr) {
					const long long v = LoadIntAt(payload, pwidth, j);
					Add128(slo, shi, (u64)v, v < 0 ? ~0ull : 0ull);
				}
				cnt++;
			}
			if (j < n && keys[j] == k) {
				bad |= 2u;
				continue;
			}
		}
		const long long sv = (long long)slo;
		const bool fits = cnt < (1ull << (64 - kDenseSumBits)) && shi == (sv < 0 ? ~0ull : 0ull) &&
		                  sv >= -(1ll << (kDenseSumBits - 1)) && sv < (1ll << (kDenseSumBits - 1));
		if (!fits) {
			bad |= 1u;
			continue;
		}
		const u64 off = (((u64)k + pack_min) ^ flip) - dmin; // sign-extended key minus the table minimum
		if (off < span) {
			dense[off] = (cnt << kDenseSumBits) | (slo & ((1ull << kDenseSumBits) - 1));
		} else {
			bad |= 1u; // cannot happen for keys inside the planned range; refuse rather than write past
		}
	}
	if (bad != 0) {
		atomicOr(flags, bad);
	}
}

} // namespace

void LaunchDenseFromRun(const void *keys, bool keys64, const void *payload, int payload_width, uint64_t n,
                        uint64_t pack_min, uint64_t flip, uint64_t dense_min, uint64_t span, uint64_t *dense,
                        unsigned int *flags_dev, hipStream_t stream) {
	FillU64(dense, 0ull, span, stream);
	FillU32(flags_dev, 0u, 1, stream);
	if (n > 0) {
		const uint64_t threads = (n + kRunItems - 1) / kRunItems;
		const uint64_t grid = (threads + 255) / 256;
		if (grid > 0x7fffffffull) {
			throw HipError("BuildDenseFromRun: too many rows for one launch");
		}
		ProfileScope prof("dense_from_run", stream, (double)n * ((keys64 ? 8 : 4) + payload_width));
		if (keys64) {
			hipLaunchKernelGGL(DenseFromRunKernel<u64>, dim3((unsigned)grid), dim3(256), 0, stream,
			                   (const u64 *)keys, payload, payload_width, (u64)n, (u64)pack_min, (u64)flip,
			                   (u64)dense_min, (u64)span, dense, flags_dev);
		} else {
			hipLaunchKernelGGL(DenseFromRunKernel<uint32_t>, dim3((unsigned)grid), dim3(256), 0, stream,
			                   (const uint32_t *)keys, payload, payload_width, (u64)n, (u64)pack_min, (u64)flip,
			                   (u64)dense_min, (u64)span, dense, flags_dev);
		}
		VGPU_HIP(hipGetLastError());
	}
}

unsigned BuildDenseFromRun(const void *keys, bool keys64, const void *payload, int payload_width, uint64_t n,
                           uint64_t pack_min, uint64_t flip, uint64_t dense_min, uint64_t span, uint64_t *dense,
                           unsigned int *flags_dev, hipStream_t stream) {
	LaunchDenseFromRun(keys, keys64, payload, payload_width, n, pack_min, flip, dense_min, span, dense, flags_dev,
	                   stream);
	unsigned int host = 0;
	VGPU_HIP(hipMemcpyAsync(&host, flags_dev, 4, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	return host;
}

const uint16_t *LaunchNarrowDense(const uint64_t *dense, uint64_t dense_span, unsigned int *nofit_dev,
                                  Scratch &scratch, hipStream_t stream) {
	if (dense == nullptr || dense_span == 0 || dense_span > (uint64_t)kPartBuckets * kSliceEntries) {
		FillU32(nofit_dev, 1u, 1, stream);
		return nullptr;
	}
	auto *dense16 = static_cast<uint16_t *>(scratch.Alloc(dense_span * 2));
	FillU32(nofit_dev, 0u, 1, stream);
	const unsigned grid = (unsigned)std::min<uint64_t>((dense_span + 255) / 256, 4096);
	hipLaunchKernelGGL(NarrowDenseKernel, dim3(grid), dim3(256), 0, stream, dense, (u64)dense_span, dense16,
	                   nofit_dev);
	VGPU_HIP(hipGetLastError());
	return dense16;
}

ProbePartition PlanProbePartition(int key_type, uint64_t n, bool has_payload, int payload_width,
                                  uint64_t dense_min, uint64_t dense_span, Scratch &scratch) {
	if (dense_span == 0 || dense_span > (1ull << 32)) {
		throw HipError("partitioned probe: needs a direct-address table of at most 2^32 entries");
	}
	if (has_payload && !(payload_width == 1 || payload_width == 2 || payload_width == 4 || payload_width == 8)) {
		throw HipError("partitioned probe: probe payload width must be 1, 2, 4 or 8 bytes");
	}
	if (key_type != VT_INT16 && key_type != VT_INT32 && key_type != VT_INT64) {
		throw HipError("partitioned probe: join keys must be INT16, INT32 or INT64");
	}
	ProbePartition p;
	p.key_type = key_type;
	p.n = n;
	p.has_payload = has_payload;
	p.payload_width = has_payload ? payload_width : 0;
	p.dense_min = dense_min;
	p.span = dense_span;
	int bits = 0;
	while (bits < 64 && (dense_span - 1) >> bits) {
		bits++;
	}
	p.shift = bits > 8 ? bits - 8 : 0; // bucket = offset >> shift < 256
	p.batch = (uint64_t)kPartBlock * (has_payload ? kScatterItemsY : kScatterItems);
	p.nbatches = (n + p.batch - 1) / p.batch;
	if (p.nbatches > 0x7fffffffull) {
		throw HipError("partitioned probe: too many probe rows for one launch");
	}
	if (n > 0) {
		p.part = static_cast<uint32_t *>(scratch.Alloc(p.nbatches * p.batch * 4));
		if (has_payload) {
			p.part_y = scratch.Alloc(p.nbatches * p.batch * (payload_width <= 4 ? 4 : 8));
		}
		p.tab = static_cast<uint32_t *>(scratch.Alloc(p.nbatches * (kPartBuckets + 1) * 4));
	}
	p.partials = static_cast<uint64_t *>(scratch.Alloc((size_t)kPartBuckets * kSliceSplits * 5 * 8));
	return p;
}

void LaunchProbePartition(const ProbePartition &p, const void *probe_keys, const void *probe_payload,
                          hipStream_t stream) {
	if (p.n == 0) {
		return;
	}
	if ((probe_payload != nullptr) != p.has_payload) {
		throw HipError("partitioned probe: probe payload does not match the plan");
	}
	switch (p.key_type) {
	case VT_INT16:
		LaunchScatter<int16_t>(p, probe_keys, probe_payload, stream);
		break;
	case VT_INT32:
		LaunchScatter<int32_t>(p, probe_keys, probe_payload, stream);
		break;
	default:
		LaunchScatter<int64_t>(p, probe_keys, probe_payload, stream);
		break;
	}
}

void LaunchPartitionedProbe(const ProbePartition &p, const uint64_t *dense, const uint16_t *dense16, uint64_t *acc,
                            hipStream_t stream) {
	if (dense == nullptr) {
		throw HipError("partitioned probe: needs a direct-address table");
	}
	FillU64(acc, 0ull, 5, stream);
	if (p.n == 0) {
		return;
	}
	if (!p.has_payload) {
		LaunchBucketProbeY<void>(p, dense, dense16, acc, stream);
	} else if (p.payload_width <= 4) {
		LaunchBucketProbeY<int32_t>(p, dense, dense16, acc, stream);
	} else {
		LaunchBucketProbeY<long long>(p, dense, dense16, acc, stream);
	}
}

bool BuildDenseIndex(int key_type, const void *keys, const uint64_t *count, const uint64_t *sum, uint64_t nd,
                     uint64_t min, uint64_t span, uint64_t *dense, unsigned int *overflow_dev, hipStream_t stream) {
	FillU64(dense, 0ull, span, stream);
	FillU32(overflow_dev, 0u, 1, stream);
	if (nd > 0) {
		const unsigned grid = (unsigned)std::min<uint64_t>((nd + 255) / 256, 4096);
		switch (key_type) {
		case VT_INT16:
			hipLaunchKernelGGL((DenseIndexKernel<int16_t>), dim3(grid), dim3(256), 0, stream, (const int16_t *)keys,
			                   count, sum, nd, min, dense, overflow_dev);
			break;
		case VT_INT32:
			hipLaunchKernelGGL((DenseIndexKernel<int32_t>), dim3(grid), dim3(256), 0, stream, (const int32_t *)keys,
			                   count, sum, nd, min, dense, overflow_dev);
			break;
		case VT_INT64:
			hipLaunchKernelGGL((DenseIndexKernel<int64_t>), dim3(grid), dim3(256), 0, stream, (const int64_t *)keys,
			                   count, sum, nd, min, dense, overflow_dev);
			break;
		default:
			throw HipError("BuildDenseIndex: join keys must be INT16, INT32 or INT64");
		}
		VGPU_HIP(hipGetLastError());
	}
	unsigned int overflow = 0;
	VGPU_HIP(hipMemcpyAsync(&overflow, overflow_dev, 4, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	return overflow == 0;
}

void RangeProbeAgg(int key_type, const void *probe_keys, const uint64_t *seg, const uint64_t *seg_row,
                   const uint64_t *seg_len, uint32_t nseg, uint64_t nvirt, const void *index_keys,
                   const uint64_t *index_count, const uint64_t *index_sum, uint64_t nd, const uint64_t *dense,
                   uint64_t dense_min, uint64_t dense_span, uint64_t *acc, hipStream_t stream) {
	const DenseArgs da {dense, dense_min, dense_span};
	FillU64(acc, 0ull, 3ull * nseg, stream);
	if (nvirt == 0) {
		return;
	}
	if (nvirt / kProbeTile >= (1ull << 31)) {
		throw HipError("RangeProbeAgg: too many probe rows for one launch");
	}
	ProfileScope prof("range_probe_agg", stream, (double)nvirt * TypeWidth(key_type));
	switch (key_type) {
	case VT_INT16:
		Launch<int16_t>(probe_keys, seg, seg_row, seg_len, nseg, nvirt, index_keys, index_count, index_sum, nd, da,
		                acc, stream);
		break;
	case VT_INT32:
		Launch<int32_t>(probe_keys, seg, seg_row, seg_len, nseg, nvirt, index_keys, index_count, index_sum, nd, da,
		                acc, stream);
		break;
	case VT_INT64:
		Launch<int64_t>(probe_keys, seg, seg_row, seg_len, nseg, nvirt, index_keys, index_count, index_sum, nd, da,
		                acc, stream);
		break;
	default:
		throw HipError("RangeProbeAgg: join keys must be INT16, INT32 or INT64");
	}
}

} // namespace vgpu
