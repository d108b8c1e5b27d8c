// join.hip -- merge-path matching of two key-sorted unique runs, pair expansion and the aggregate combine
// of the sort-merge equi-join (gfx950).
//
// Replaces the reference's right-side sort + per-left-row lower_bound/upper_bound + exclusive_scan +
// one-thread-per-left-row expansion (cuda_engine/src/operators/gpu_hash_join.cu:196-254). Both sides are
// first reduced to their distinct keys (radix sort + segmented reduction, with per-key row counts and any
// aggregate inputs); this file then
//   * partitions the merge of the two sorted key runs into equal tiles (merge path: one diagonal search
//     per tile), stages the tile's B segment in LDS and finds each A key's partner with an LDS binary
//     search -- every key is read once, no random global probes;
//   * for an aggregate-only join, folds COUNT(*) = sum ca*cb and the per-side sums/min/max of the matched
//     keys in exact int128 / double-double, per block, then in one deterministic final block;
//   * for a materialized join, scans the per-key pair counts and expands every output row from its own
//     index (binary search over the scanned offsets), so a key with millions of matches is spread across
//     the whole device instead of one thread.
#include "agg_state.h"
#include "common.h"
#include "kernels.h"
#include "join_kernels.h"

#include <algorithm>

namespace vgpu {
namespace {

constexpr int kMergeBlock = 256;
constexpr int kMergeTile = 2048; // merged items per tile

__device__ __forceinline__ u64 KeyAt(const void *keys, uint64_t i, int width, bool is_signed) {
	return OrderedKey(LoadRawBits(keys, i, width), width, is_signed);
}

//! Merge-path split of diagonal d: the number of A items among the first d merged items, with A before B
//! on equal keys.
__device__ __forceinline__ uint64_t MergePathSplit(const void *a, uint64_t na, const void *b, uint64_t nb, uint64_t d,
                                                   int width, bool is_signed) {
	uint64_t lo = d > nb ? d - nb : 0;
	uint64_t hi = d < na ? d : na;
	while (lo < hi) {
		uint64_t mid = (lo + hi) >> 1; // candidate: mid A items, d - mid B items
		// A[mid] must come after B[d - mid - 1]: take more A while A[mid] <= B[d - mid - 1]
		if (KeyAt(a, mid, width, is_signed) <= KeyAt(b, d - mid - 1, width, is_signed)) {
			lo = mid + 1;
		} else {
			hi = mid;
		}
	}
	return lo;
}

__global__ void __launch_bounds__(kMergeBlock) MergeMatchKernel(const void *akeys, uint64_t na, const void *bkeys,
                                                                uint64_t nb, int width, int is_signed,
                                                                long long *match) {
	__shared__ u64 s_b[kMergeTile + 1];
	__shared__ uint64_t s_split[2];
	const uint64_t d0 = (uint64_t)blockIdx.x * kMergeTile;
	const uint64_t total = na + nb;
	const uint64_t d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
	if (threadIdx.x == 0) {
		s_split[0] = MergePathSplit(akeys, na, bkeys, nb, d0, width, is_signed != 0);
	} else if (threadIdx.x == 64) {
		s_split[1] = MergePathSplit(akeys, na, bkeys, nb, d1, width, is_signed != 0);
	}
	__syncthreads();
	const uint64_t i0 = s_split[0], i1 = s_split[1];
	const uint64_t j0 = d0 - i0, j1 = d1 - i1;
	const uint32_t bn = (uint32_t)(j1 - j0);
	// B segment plus the first B item of the next tile (an A key's equal partner may sit there).
	for (uint32_t t = threadIdx.x; t <= bn; t += kMergeBlock) {
		uint64_t j = j0 + t;
		s_b[t] = j < nb ? KeyAt(bkeys, j, width, is_signed != 0) : ~0ull;
	}
	__syncthreads();
	for (uint64_t i = i0 + threadIdx.x; i < i1; i += kMergeBlock) {
		u64 k = KeyAt(akeys, i, width, is_signed != 0);
		// lower_bound within [0, bn]
		uint32_t lo = 0, hi = bn;
		while (lo < hi) {
			uint32_t mid = (lo + hi) >> 1;
			if (s_b[mid] < k) {
				lo = mid + 1;
			} else {
				hi = mid;
			}
		}
		uint64_t j = j0 + lo;
		match[i] = (j < nb && s_b[lo] == k) ? (long long)j : -1ll;
	}
}

//! pairs[i] = ca[i] * cb[match[i]] (0 when unmatched).
__global__ void PairCountKernel(const long long *match, const long long *ca, const long long *cb, uint64_t na,
                                u64 *pairs) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += stride) {
		long long m = match[i];
		pairs[i] = m >= 0 ? (u64)ca[i] * (u64)cb[m] : 0ull;
	}
}

// ---------------------------------------------------------------- exclusive scan (u64, look-back)

constexpr int kScanBlock = 256;
constexpr int kScanItems = 32; // 8,192 values per tile: one ticket (a serialized same-address atomic) per tile

__global__ void __launch_bounds__(kScanBlock) ScanU64Kernel(const u64 *in, u64 *out, uint64_t n, u64 *state,
                                                            uint32_t *ticket, u64 *total_out, uint64_t ntiles) {
	__shared__ uint32_t s_tile;
	__shared__ u64 s_scan[kScanBlock / kWave + 1];
	__shared__ u64 s_prefix;
	if (threadIdx.x == 0) {
		s_tile = FetchAddAgent(ticket, 1u);
	}
	__syncthreads();
	const uint32_t tile = s_tile;
	const uint64_t base = (uint64_t)tile * kScanBlock * kScanItems + (uint64_t)threadIdx.x * kScanItems;
	u64 v[kScanItems];
	u64 sum = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		v[k] = base + k < n ? in[base + k] : 0ull;
		sum += v[k];
	}
	u64 total;
	u64 excl = BlockExclusiveScan<u64, kScanBlock>(sum, s_scan, &total);
	if (threadIdx.x == 0) {
		StoreAgent(state + tile, LbPack(tile == 0 ? kLbInclusive : kLbAggregate, total));
	}
	if (threadIdx.x < kWave) {
		u64 pre = tile == 0 ? 0ull : LookbackExclusiveWave(state, tile);
		if (threadIdx.x == 0) {
			if (tile > 0) {
				StoreAgent(state + tile, LbPack(kLbInclusive, pre + total));
			}
			s_prefix = pre;
			if (tile == ntiles - 1 && total_out != nullptr) {
				*total_out = pre + total;
			}
		}
	}
	__syncthreads();
	u64 run = s_prefix + excl;
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		if (base + k < n) {
			out[base + k] = run;
			run += v[k];
		}
	}
}

//! Output row o belongs to A-key m = upper_bound(offsets, o) - 1; within that key's ca x cb block it is
//! (l, r) = (q / cb, q % cb) with q = o - offsets[m]. Writes the two source row ids.
__global__ void ExpandPairsKernel(const u64 *offsets, uint64_t na, const long long *match, const long long *cb,
                                  const long long *starta, const long long *startb, const uint32_t *perma,
                                  const uint32_t *permb, uint64_t total, uint32_t *out_left, uint32_t *out_right) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += stride) {
		uint64_t lo = 0, hi = na;
		while (lo < hi) {
			uint64_t mid = (lo + hi) >> 1;
			if (offsets[mid] <= o) {
				lo = mid + 1;
			} else {
				hi = mid;
			}
		}
		uint64_t m = lo - 1;
		u64 q = o - offsets[m];
		long long mb = match[m];
		u64 nbk = (u64)cb[mb];
		u64 l = q / nbk, r = q - l * nbk;
		out_left[o] = perma[(u64)starta[m] + l];
		out_right[o] = permb[(u64)startb[mb] + r];
	}
}

// ---------------------------------------------------------------- aggregate combine

constexpr int kAggBlock = 256;

__device__ __forceinline__ void TwoProd(double a, double b, double &p, double &e) {
	p = a * b;
	e = __fma_rn(a, b, -p); // error-free product (explicit fused multiply-add)
}

//! Per-block partial of every join aggregate over matched keys; state layout = JoinAggState.
__global__ void __launch_bounds__(kAggBlock) JoinAggKernel(JoinAggArgs args, u64 *partials) {
	__shared__ u64 s_part[kAggBlock / kWave][kJoinAggMax * 2];
	u64 acc[kJoinAggMax * 2];
#pragma unroll
	for (int q = 0; q < kJoinAggMax; q++) {
		JoinAggIdentity(args, q, acc[2 * q], acc[2 * q + 1]);
	}
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < args.na; i += stride) {
		long long m = args.match[i];
		if (m < 0) {
			continue;
		}
		JoinAggAccumulate(args, i, (uint64_t)m, acc);
	}
	// wave reduce in lane order, then waves in order
	const int lane = LaneId(), wave = threadIdx.x / kWave;
	for (int off = 1; off < kWave; off <<= 1) {
		u64 other[kJoinAggMax * 2];
#pragma unroll
		for (int s = 0; s < kJoinAggMax * 2; s++) {
			other[s] = __shfl_xor(acc[s], off);
		}
		if ((lane & off) == 0) {
			JoinAggMerge(args, acc, other);
		} else {
			JoinAggMerge(args, other, acc);
#pragma unroll
			for (int s = 0; s < kJoinAggMax * 2; s++) {
				acc[s] = other[s];
			}
		}
	}
	if (lane == 0) {
#pragma unroll
		for (int s = 0; s < kJoinAggMax * 2; s++) {
			s_part[wave][s] = acc[s];
		}
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		u64 tot[kJoinAggMax * 2];
#pragma unroll
		for (int s = 0; s < kJoinAggMax * 2; s++) {
			tot[s] = s_part[0][s];
		}
		for (int w = 1; w < kAggBlock / kWave; w++) {
			u64 t[kJoinAggMax * 2];
#pragma unroll
			for (int s = 0; s < kJoinAggMax * 2; s++) {
				t[s] = s_part[w][s];
			}
			JoinAggMerge(args, tot, t);
		}
#pragma unroll
		for (int s = 0; s < kJoinAggMax * 2; s++) {
			partials[(uint64_t)blockIdx.x * kJoinAggMax * 2 + s] = tot[s];
		}
	}
}

//! One wave folds the per-block partials: lane l merges partials l, l + 64, ... in order, then the lanes
//! merge in lane order (a fixed tree, so the double-double sums round the same way on every run).
__global__ void __launch_bounds__(kWave) JoinAggFinalKernel(JoinAggArgs args, const u64 *partials, uint32_t nparts,
                                                            u64 *result) {
	const int lane = LaneId();
	u64 tot[kJoinAggMax * 2];
#pragma unroll
	for (int q = 0; q < kJoinAggMax; q++) {
		JoinAggIdentity(args, q, tot[2 * q], tot[2 * q + 1]);
	}
	for (uint32_t p = lane; p < nparts; p += kWave) {
		u64 t[kJoinAggMax * 2];
#pragma unroll
		for (int s = 0; s < kJoinAggMax * 2; s++) {
			t[s] = partials[(uint64_t)p * kJoinAggMax * 2 + s];
		}
		JoinAggMerge(args, tot, t);
	}
	for (int off = 1; off < kWave; off <<= 1) {
		u64 other[kJoinAggMax * 2];
#pragma unroll
		for (int s = 0; s < kJoinAggMax * 2; s++) {
			other[s] = __shfl_xor(tot[s], off);
		}
		if ((lane & off) == 0) {
			JoinAggMerge(args, tot, other);
		} else {
			JoinAggMerge(args, other, tot);
#pragma unroll
			for (int s = 0; s < kJoinAggMax * 2; s++) {
				tot[s] = other[s];
			}
		}
	}
	if (lane == 0) {
#pragma unroll
		for (int s = 0; s < kJoinAggMax * 2; s++) {
			result[s] = tot[s];
		}
	}
}

unsigned Grid(uint64_t n, unsigned cap = 8192) {
	uint64_t g = (n + 255) / 256;
	return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

} // namespace

void MergeMatch(const void *akeys, uint64_t na, const void *bkeys, uint64_t nb, int vtype, long long *match,
                hipStream_t stream) {
	if (na == 0) {
		return;
	}
	const uint64_t tiles = (na + nb + kMergeTile - 1) / kMergeTile;
	ProfileScope prof("merge_match", stream, (double)(na + nb) * TypeWidth(vtype) + (double)na * 8);
	hipLaunchKernelGGL(MergeMatchKernel, dim3((unsigned)tiles), dim3(kMergeBlock), 0, stream, akeys, na, bkeys, nb,
	                   TypeWidth(vtype), TypeIsSigned(vtype) ? 1 : 0, match);
	VGPU_HIP(hipGetLastError());
}

void PairCounts(const long long *match, const long long *ca, const long long *cb, uint64_t na, uint64_t *pairs,
                hipStream_t stream) {
	if (na == 0) {
		return;
	}
	hipLaunchKernelGGL(PairCountKernel, dim3(Grid(na)), dim3(256), 0, stream, match, ca, cb, na, (u64 *)pairs);
	VGPU_HIP(hipGetLastError());
}

void ExclusiveScanU64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_device, Scratch &scratch,
                      hipStream_t stream) {
	if (n == 0) {
		if (total_device != nullptr) {
			FillU64(total_device, 0ull, 1, stream);
		}
		return;
	}
	const uint64_t per = (uint64_t)kScanBlock * kScanItems;
	const uint64_t tiles = (n + per - 1) / per;
	auto *state = static_cast<u64 *>(scratch.Alloc((tiles + 1) * sizeof(u64)));
	FillU64(state, 0ull, tiles + 1, stream);
	hipLaunchKernelGGL(ScanU64Kernel, dim3((unsigned)tiles), dim3(kScanBlock), 0, stream, (const u64 *)in, (u64 *)out,
	                   n, state, reinterpret_cast<uint32_t *>(state + tiles), (u64 *)total_device, tiles);
	VGPU_HIP(hipGetLastError());
}

void ExpandPairs(const uint64_t *offsets, uint64_t na, const long long *match, const long long *cb,
                 const long long *starta, const long long *startb, const uint32_t *perma, const uint32_t *permb,
                 uint64_t total, uint32_t *out_left, uint32_t *out_right, hipStream_t stream) {
	if (total == 0) {
		return;
	}
	ProfileScope prof("join_expand", stream, (double)total * 8);
	hipLaunchKernelGGL(ExpandPairsKernel, dim3(Grid(total, 16384)), dim3(256), 0, stream, (const u64 *)offsets, na,
	                   match, cb, starta, startb, perma, permb, total, out_left, out_right);
	VGPU_HIP(hipGetLastError());
}

void JoinAggregate(const JoinAggArgs &args, uint64_t *result_device, Scratch &scratch, hipStream_t stream) {
	const unsigned grid = Grid(args.na, 1024);
	auto *partials = static_cast<u64 *>(scratch.Alloc((size_t)grid * kJoinAggMax * 2 * sizeof(u64)));
	{
		ProfileScope prof("join_agg", stream, (double)args.na * 16);
		hipLaunchKernelGGL(JoinAggKernel, dim3(grid), dim3(kAggBlock), 0, stream, args, partials);
		VGPU_HIP(hipGetLastError());
	}
	hipLaunchKernelGGL(JoinAggFinalKernel, dim3(1), dim3(kWave), 0, stream, args, partials, grid, (u64 *)result_device);
	VGPU_HIP(hipGetLastError());
}

} // namespace vgpu
