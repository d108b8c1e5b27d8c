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

} // namespace

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
