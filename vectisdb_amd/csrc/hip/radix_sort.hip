// radix_sort.hip -- LSD radix sort for gfx950: key-range pre-pass, one-read multi-digit histogram, and a
// onesweep digit pass (LDS-staged, wave64 multisplit ranking, decoupled look-back on per-digit counts).
//
// Replaces the reference's `thrust::sequence` + `thrust::stable_sort_by_key` per key column
// (cuda_engine/src/operators/gpu_groupby.cu:352-376) and the right-side `thrust::sort_by_key` of the join
// (cuda_engine/src/operators/gpu_hash_join.cu:205-206). Differences by design:
//   * the key is the order-preserving packing of ALL key columns, minus their minimum, so only the bits
//     that vary are sorted (a 20-bit key range takes 2-3 digit passes, not 8);
//   * payload columns (the aggregate inputs) travel with the key through LDS, so no per-aggregate gather
//     through a permutation is needed afterwards;
//   * one kernel per digit pass: each tile ranks its keys in LDS, publishes its per-digit counts, and
//     learns its global offsets by looking back over predecessor tiles (agent-scope atomics, visible
//     across all 8 XCDs), so keys are read once and written once per pass.
#include "common.h"
#include "kernels.h"
#include "onesweep.h"

#include <algorithm>
#include <cstring>

namespace vgpu {
namespace {

using namespace sortpass;

struct RangeCols {
	int ncols;
	const void *col[kMaxKeyCols];
	int width[kMaxKeyCols];
	int is_signed[kMaxKeyCols];
};

__global__ void __launch_bounds__(256) KeyRangeKernel(RangeCols cols, uint64_t n, u64 *mins, u64 *maxs) {
	u64 lo[kMaxKeyCols], hi[kMaxKeyCols];
#pragma unroll
	for (int c = 0; c < kMaxKeyCols; c++) {
		lo[c] = ~0ull;
		hi[c] = 0;
	}
	// Four strided rows per step so each thread keeps four loads in flight.
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t row0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; row0 < n; row0 += 4 * stride) {
#pragma unroll
		for (int c = 0; c < kMaxKeyCols; c++) {
			if (c < cols.ncols) {
				u64 raw[4];
#pragma unroll
				for (int u = 0; u < 4; u++) {
					uint64_t row = row0 + u * stride;
					raw[u] = row < n ? LoadRawBits(cols.col[c], row, cols.width[c]) : 0ull;
				}
#pragma unroll
				for (int u = 0; u < 4; u++) {
					if (row0 + u * stride < n) {
						u64 ord = OrderedKey(raw[u], cols.width[c], cols.is_signed[c] != 0);
						lo[c] = ord < lo[c] ? ord : lo[c];
						hi[c] = ord > hi[c] ? ord : hi[c];
					}
				}
			}
		}
	}
#pragma unroll
	for (int c = 0; c < kMaxKeyCols; c++) {
		if (c < cols.ncols) {
			for (int off = 32; off >= 1; off >>= 1) {
				u64 a = __shfl_xor(lo[c], off);
				u64 b = __shfl_xor(hi[c], off);
				lo[c] = a < lo[c] ? a : lo[c];
				hi[c] = b > hi[c] ? b : hi[c];
			}
			if (LaneId() == 0) {
				MinAgentIfLower(mins + c, lo[c]);
				MaxAgentIfHigher(maxs + c, hi[c]);
			}
		}
	}
}

//! Loads in flight per lane per loop step of the key-range pass (enough bytes in flight to cover HBM
//! latency with one wave generation on the chip).
constexpr int kHistUnroll = 8;

//! Counters per histogram bin in the key-range pass (spread over lanes; see KeyRangeHist8Kernel).
constexpr int kHistCopies = 8;

//! Min/max of the ordered key and the histograms of all 8 byte digits in one read. A digit that is the
//! same for every key a wave reads in one step (the constant high bytes of a narrow key range) is counted
//! with one LDS atomic instead of one per key. Each bin has kHistCopies adjacent counters, lane l adding to
//! copy l % kHistCopies: lanes of one wave that add to the same bin (equal keys of a skewed distribution, or
//! a byte with few values) then mostly hit different banks instead of serializing on one word.
template <typename T, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8)))
KeyRangeHist8Kernel(const T *col, int width, int is_signed, uint64_t n, u64 *mins, u64 *maxs, uint32_t *hist8) {
	__shared__ uint32_t s_hist[8 * 256 * kHistCopies];
	for (int i = threadIdx.x; i < 8 * 256 * kHistCopies; i += blockDim.x) {
		s_hist[i] = 0;
	}
	__syncthreads();
	const int lane = LaneId();
	const uint32_t copy = (uint32_t)lane & (kHistCopies - 1);
	u64 lo = ~0ull, hi = 0;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	const uint64_t start = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const uint64_t wave_start = start - lane;
	for (uint64_t row0 = start, w0 = wave_start; w0 < n; row0 += kHistUnroll * stride, w0 += kHistUnroll * stride) {
		u64 raw[kHistUnroll];
#pragma unroll
		for (int u = 0; u < kHistUnroll; u++) {
			uint64_t row = row0 + u * stride;
			raw[u] = row < n ? (u64)col[row] : 0ull;
		}
		// Digits that are the same for every key of this step (the constant high bytes of a narrow key
		// range) are found once per step: each lane ORs its keys' differences from one reference key, and a
		// wave OR tells which bytes vary anywhere. A constant byte is counted with one atomic for all the
		// step's keys; only the varying bytes are counted key by key.
		u64 ord[kHistUnroll];
		uint32_t nvalid = 0;
#pragma unroll
		for (int u = 0; u < kHistUnroll; u++) {
			ord[u] = OrderedKey(raw[u], width, is_signed != 0);
			nvalid += (uint32_t)Popc64(__ballot(row0 + u * stride < n));
		}
		const u64 vmask0 = __ballot(row0 < n); // item 0 has the lowest rows: empty only if the step is empty
		if (vmask0 == 0) {
			continue;
		}
		const int leader = __ffsll((long long)vmask0) - 1;
		const u64 ref = ((u64)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(ord[0] >> 32), leader) << 32) |
		                (u64)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)ord[0], leader);
		u64 diff = 0;
#pragma unroll
		for (int u = 0; u < kHistUnroll; u++) {
			if (row0 + u * stride < n) {
				diff |= ord[u] ^ ref;
				lo = ord[u] < lo ? ord[u] : lo;
				hi = ord[u] > hi ? ord[u] : hi;
			}
		}
		uint32_t bytes = 0; // bit d: byte d differs from the reference in some key of this lane
#pragma unroll
		for (int d = 0; d < 8; d++) {
			bytes |= ((diff >> (8 * d)) & 255u) != 0 ? (1u << d) : 0u;
		}
		for (int off = 32; off >= 1; off >>= 1) {
			bytes |= (uint32_t)__shfl_xor((int)bytes, off);
		}
		const uint32_t vary = (uint32_t)__builtin_amdgcn_readfirstlane((int)bytes);
#pragma unroll
		for (int d = 0; d < 8; d++) {
			if (((vary >> d) & 1u) == 0) {
				if (lane == leader) {
					atomicAdd(&s_hist[(d * 256 + ((uint32_t)(ref >> (8 * d)) & 255u)) * kHistCopies], nvalid);
				}
			} else {
#pragma unroll
				for (int u = 0; u < kHistUnroll; u++) {
					if (row0 + u * stride < n) {
						atomicAdd(&s_hist[(d * 256 + ((uint32_t)(ord[u] >> (8 * d)) & 255u)) * kHistCopies + copy], 1u);
					}
				}
			}
		}
	}
	__syncthreads();
	for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) {
		uint32_t v = 0;
#pragma unroll
		for (int c = 0; c < kHistCopies; c++) {
			v += s_hist[i * kHistCopies + c];
		}
		if (v) {
			FetchAddAgent(hist8 + i, v);
		}
	}
	for (int off = 32; off >= 1; off >>= 1) {
		u64 a = __shfl_xor(lo, off);
		u64 b = __shfl_xor(hi, off);
		lo = a < lo ? a : lo;
		hi = b > hi ? b : hi;
	}
	if (lane == 0) {
		MinAgentIfLower(mins, lo);
		MaxAgentIfHigher(maxs, hi);
	}
}

//! Clears the byte-digit histograms and sets the min/max words (min all ones, max zero) in one launch.
__global__ void __launch_bounds__(1024) KeyRangeInitKernel(uint32_t *hist8, u64 *mm) {
	for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) {
		hist8[i] = 0;
	}
	if (threadIdx.x == 0) {
		mm[0] = ~0ull;
		mm[1] = 0;
	}
}

//! One read of the keys builds the digit histogram of every pass.
template <int R>
__global__ void __launch_bounds__(256) HistKernel(KeyPack pack, uint64_t n, PassList passes, uint32_t *hist,
                                                  const uint64_t *n_dev) {
	constexpr int RADIX = 1 << R;
	if (n_dev != nullptr) {
		n = *n_dev; // row count written by an earlier kernel (n was its bound)
	}
	constexpr int PMAX = (64 + R - 1) / R;
	constexpr int C = R == 8 ? 4 : 1; // lane copies per bin (as in KeyRangeHist8Kernel) where LDS allows
	__shared__ uint32_t s_hist[PMAX * RADIX * C];
	for (int i = threadIdx.x; i < PMAX * RADIX * C; i += blockDim.x) {
		s_hist[i] = 0;
	}
	__syncthreads();
	const uint32_t copy = (uint32_t)LaneId() & (C - 1);
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t row0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; row0 < n; row0 += 4 * stride) {
		u64 key[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			uint64_t row = row0 + u * stride;
			key[u] = row < n ? PackRow(pack, row) : 0ull;
		}
#pragma unroll
		for (int u = 0; u < 4; u++) {
			if (row0 + u * stride < n) {
				for (int p = 0; p < passes.npasses; p++) {
					uint32_t d = (uint32_t)(key[u] >> passes.shift[p]) & ((1u << passes.bits[p]) - 1);
					atomicAdd(&s_hist[(p * RADIX + d) * C + copy], 1u);
				}
			}
		}
	}
	__syncthreads();
	for (int i = threadIdx.x; i < passes.npasses * RADIX; i += blockDim.x) {
		uint32_t v = 0;
#pragma unroll
		for (int c = 0; c < C; c++) {
			v += s_hist[i * C + c];
		}
		if (v) {
			FetchAddAgent(hist + i, v);
		}
	}
}

//! Keys packed from several columns: one read of the key columns writes the packed keys (so the first radix
//! pass reads one compact key instead of every key column, one column at a time) and builds the 8-bit digit
//! histogram of every pass. A digit that a whole wave shares (a composite key's leading column, constant over
//! long runs of rows) is counted by one lane with one LDS add instead of 64 colliding ones; each wave counts
//! into its own copy of the histograms, summed at the end (lane copies in the same LDS measured 3% slower).
template <typename KeyT>
__global__ void __launch_bounds__(256) PackHistKernel(KeyPack pack, uint64_t n, PassList passes, KeyT *keys_out,
                                                      uint32_t *hist, const uint64_t *n_dev) {
	constexpr int RADIX = 256;
	constexpr int PMAX = 8;
	constexpr int kWaves = 256 / kWave;
	__shared__ uint32_t s_hist[kWaves][PMAX * RADIX];
	if (n_dev != nullptr) {
		n = *n_dev; // row count written by an earlier kernel (n was its bound)
	}
	for (int i = threadIdx.x; i < kWaves * PMAX * RADIX; i += blockDim.x) {
		(&s_hist[0][0])[i] = 0;
	}
	__syncthreads();
	const int lane = LaneId();
	uint32_t *mine = s_hist[threadIdx.x / kWave];
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	// every lane of a wave runs the same trip count (the ballots need the whole wave); four rows per lane are
	// loaded before any is counted, to keep enough reads in flight
	const uint64_t wave_row0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1));
	for (uint64_t base = wave_row0; base < n; base += 4 * stride) {
		u64 key[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const uint64_t row = base + u * stride + lane;
			key[u] = row < n ? PackRow(pack, row) : 0ull;
		}
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const uint64_t row = base + u * stride + lane;
			const bool valid = row < n;
			if (valid) {
				keys_out[row] = (KeyT)key[u];
			}
			const u64 live = __ballot(valid);
			for (int p = 0; p < passes.npasses; p++) {
				const uint32_t d = (uint32_t)(key[u] >> passes.shift[p]) & ((1u << passes.bits[p]) - 1);
				const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
				if (__ballot(valid && d == d0) == live) {
					// the whole wave holds one digit: one add instead of 64 to one address
					if (lane == 0) {
						atomicAdd(&mine[p * RADIX + d0], (uint32_t)Popc64(live));
					}
				} else if (valid) {
					atomicAdd(&mine[p * RADIX + d], 1u);
				}
			}
		}
	}
	__syncthreads();
	for (int i = threadIdx.x; i < passes.npasses * RADIX; i += blockDim.x) {
		uint32_t v = 0;
#pragma unroll
		for (int w = 0; w < kWaves; w++) {
			v += s_hist[w][i];
		}
		if (v) {
			FetchAddAgent(hist + i, v);
		}
	}
}

//! Items per thread of the kernel variant that runs (11-bit digits always use 16).
int EffectiveItems(int R, int items) {
	return (R == 8 && items == 32) ? 32 : 16;
}

//! Threads per tile of the kernel variant that runs: 512 only for 8-bit digits at 16 items without the
//! payload prefetch (the one 512-thread variant built), else 256.
int EffectiveBlock(int R, const SortTuning &t) {
	if (R == 8 && t.block >= 512 && EffectiveItems(R, t.items) == 16 && t.prefetch == 0) {
		return t.block >= 1024 ? 1024 : 512;
	}
	return kSortBlock;
}

//! One digit pass with the variant the tuning selects: 8-bit digits come with 16 or 32 items per thread and
//! with or without the payload prefetch, or as 512 threads x 16 items without it; 11-bit digits with 16
//! items and the prefetch.
void LaunchPass(int R, bool keys64, bool first, const KeyPack &pack, const void *keys_in, void *keys_out,
                const PayloadArgs &pay, uint64_t n, int shift, int bits, const uint32_t *gh, uint32_t *state,
                uint32_t *ticket, hipStream_t stream, const SortTuning &tuning, const uint64_t *n_dev) {
	const int it = EffectiveItems(R, tuning.items);
	const int block = EffectiveBlock(R, tuning);
	const bool prefetch = tuning.prefetch != 0;
	const uint64_t tile_rows = (uint64_t)block * it;
	const unsigned tiles = (unsigned)((n + tile_rows - 1) / tile_rows);
	const size_t kw = keys64 ? 8 : 4;
	double key_in = first ? 0.0 : (double)kw;
	if (first) {
		for (int c = 0; c < pack.ncols; c++) {
			key_in += pack.width[c];
		}
	}
	double pay_bytes = 0;
	for (int p = 0; p < pay.n; p++) {
		const int in_width = (first && ((pay.narrow >> p) & 1u) != 0) ? 8 : pay.width[p];
		pay_bytes += (first && pay.in[p] == nullptr ? 0.0 : in_width) + pay.width[p];
	}
	ProfileScope prof(first ? "radix_pass_first" : "radix_pass", stream, (double)n * (key_in + kw + pay_bytes));
	if (R == 11) {
		LaunchVariant<11, 16, true, 256>(first, keys64, pack, keys_in, keys_out, pay, n, shift, bits, gh, state,
		                                 ticket, stream, tiles, n_dev);
	} else if (block == 1024) {
		LaunchVariant<8, 16, false, 1024>(first, keys64, pack, keys_in, keys_out, pay, n, shift, bits, gh, state,
		                                  ticket, stream, tiles, n_dev);
	} else if (block == 512) {
		LaunchVariant<8, 16, false, 512>(first, keys64, pack, keys_in, keys_out, pay, n, shift, bits, gh, state,
		                                 ticket, stream, tiles, n_dev);
	} else if (it == 32) {
		if (prefetch) {
			LaunchVariant<8, 32, true, 256>(first, keys64, pack, keys_in, keys_out, pay, n, shift, bits, gh, state,
			                                ticket, stream, tiles, n_dev);
		} else {
			LaunchVariant<8, 32, false, 256>(first, keys64, pack, keys_in, keys_out, pay, n, shift, bits, gh, state,
			                                 ticket, stream, tiles, n_dev);
		}
	} else if (prefetch) {
		LaunchVariant<8, 16, true, 256>(first, keys64, pack, keys_in, keys_out, pay, n, shift, bits, gh, state,
		                                ticket, stream, tiles, n_dev);
	} else {
		LaunchVariant<8, 16, false, 256>(first, keys64, pack, keys_in, keys_out, pay, n, shift, bits, gh, state,
		                                 ticket, stream, tiles, n_dev);
	}
}

} // namespace

void KeyRanges(const void *const *cols, const int *vtypes, int ncols, uint64_t n, uint64_t *min_out,
               uint64_t *max_out, Scratch &scratch, hipStream_t stream) {
	if (ncols <= 0 || ncols > kMaxKeyCols) {
		throw HipError("KeyRanges: between 1 and 4 key columns are supported");
	}
	RangeCols rc {};
	rc.ncols = ncols;
	for (int c = 0; c < ncols; c++) {
		rc.col[c] = cols[c];
		rc.width[c] = TypeWidth(vtypes[c]);
		rc.is_signed[c] = TypeIsSigned(vtypes[c]) ? 1 : 0;
	}
	auto *dev = static_cast<u64 *>(scratch.Alloc(2 * kMaxKeyCols * sizeof(u64)));
	FillU64(dev, ~0ull, kMaxKeyCols, stream);
	FillU64(dev + kMaxKeyCols, 0ull, kMaxKeyCols, stream);
	if (n > 0) {
		uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
		double bytes = 0;
		for (int c = 0; c < ncols; c++) {
			bytes += (double)n * rc.width[c];
		}
		ProfileScope prof("key_range", stream, bytes);
		hipLaunchKernelGGL(KeyRangeKernel, dim3((unsigned)blocks), dim3(256), 0, stream, rc, n, dev,
		                   dev + kMaxKeyCols);
		VGPU_HIP(hipGetLastError());
	}
	u64 host[2 * kMaxKeyCols];
	ReadBack(host, dev, sizeof(host), stream);
	for (int c = 0; c < ncols; c++) {
		min_out[c] = host[c];
		max_out[c] = host[kMaxKeyCols + c];
	}
}

void KeyRangeHist8(const void *col, int vtype, uint64_t n, uint64_t *min_out, uint64_t *max_out,
                   uint32_t *hist8_dev, hipStream_t stream) {
	// hist8_dev is followed by 2 words of min/max scratch (callers allocate 8 * 256 + 4 uint32).
	u64 *mm = reinterpret_cast<u64 *>(hist8_dev + 8 * 256);
	hipLaunchKernelGGL(KeyRangeInitKernel, dim3(1), dim3(1024), 0, stream, hist8_dev, mm);
	VGPU_HIP(hipGetLastError());
	if (n > 0) {
		const int width = TypeWidth(vtype);
		// Every block merges its LDS histograms into the global ones with device-scope atomics at its end, so
		// the block count is kept low: 1,024-thread blocks, at most 512 (two per CU, 32 waves per CU).
		constexpr int kHB = 1024;
		uint64_t blocks = std::min<uint64_t>((n + 4 * kHB - 1) / (4 * kHB), 512);
		ProfileScope prof("key_range_hist", stream, (double)n * width);
		const int sg = TypeIsSigned(vtype) ? 1 : 0;
		switch (width) {
		case 8:
			hipLaunchKernelGGL((KeyRangeHist8Kernel<u64, kHB>), dim3((unsigned)blocks), dim3(kHB), 0, stream,
			                   (const u64 *)col, width, sg, n, mm, mm + 1, hist8_dev);
			break;
		case 4:
			hipLaunchKernelGGL((KeyRangeHist8Kernel<uint32_t, kHB>), dim3((unsigned)blocks), dim3(kHB), 0, stream,
			                   (const uint32_t *)col, width, sg, n, mm, mm + 1, hist8_dev);
			break;
		case 2:
			hipLaunchKernelGGL((KeyRangeHist8Kernel<uint16_t, kHB>), dim3((unsigned)blocks), dim3(kHB), 0, stream,
			                   (const uint16_t *)col, width, sg, n, mm, mm + 1, hist8_dev);
			break;
		default:
			hipLaunchKernelGGL((KeyRangeHist8Kernel<uint8_t, kHB>), dim3((unsigned)blocks), dim3(kHB), 0, stream,
			                   (const uint8_t *)col, width, sg, n, mm, mm + 1, hist8_dev);
			break;
		}
		VGPU_HIP(hipGetLastError());
	}
	u64 host[2];
	ReadBack(host, mm, sizeof(host), stream);
	*min_out = host[0];
	*max_out = host[1];
}

void PlanKeyPackAligned(const void *col, int vtype, uint64_t omin, uint64_t omax, KeyPack &pack) {
	pack = KeyPack {};
	pack.ncols = 1;
	pack.col[0] = col;
	pack.width[0] = TypeWidth(vtype);
	pack.is_signed[0] = TypeIsSigned(vtype) ? 1 : 0;
	uint64_t diff = omin ^ omax;
	int span = 0;
	while (span < 64 && (diff >> span) != 0) {
		span++;
	}
	int bits = ((span + 7) / 8) * 8;
	if (bits == 0) {
		bits = 8; // one (trivial) pass keeps the sort path uniform
	}
	pack.bits[0] = bits;
	pack.shift[0] = 0;
	pack.total_bits = bits;
	pack.min[0] = bits >= 64 ? 0ull : (omin & ~((1ull << bits) - 1));
	pack.max_packed = omax - pack.min[0];
}

bool PlanKeyPack(const void *const *cols, const int *vtypes, int ncols, const uint64_t *mins, const uint64_t *maxs,
                 KeyPack &pack) {
	pack = KeyPack {};
	pack.ncols = ncols;
	int total = 0;
	for (int c = 0; c < ncols; c++) {
		pack.col[c] = cols[c];
		pack.width[c] = TypeWidth(vtypes[c]);
		pack.is_signed[c] = TypeIsSigned(vtypes[c]) ? 1 : 0;
		pack.min[c] = mins[c];
		uint64_t range = maxs[c] >= mins[c] ? maxs[c] - mins[c] : 0;
		int bits = 0;
		while (bits < 64 && (range >> bits) != 0) {
			bits++;
		}
		pack.bits[c] = bits;
		total += bits;
	}
	if (total > 64) {
		return false;
	}
	pack.total_bits = total;
	int shift = 0;
	for (int c = ncols - 1; c >= 0; c--) {
		pack.shift[c] = pack.bits[c] > 0 ? shift : 0;
		shift += pack.bits[c];
	}
	for (int c = 0; c < ncols; c++) {
		if (pack.bits[c] > 0) {
			pack.max_packed |= (maxs[c] - mins[c]) << pack.shift[c];
		}
	}
	return true;
}

void SortPackedWith(const KeyPack &pack, const void *const *payload_in, const int *payload_width, int npayload,
                    uint64_t n, const SortTuning &tuning, SortedRun &out, Scratch &scratch, hipStream_t stream,
                    const uint64_t *n_dev) {
	if (n >= (1ull << 30)) {
		throw HipError("radix sort: at most 2^30 - 1 rows per sort call");
	}
	if (npayload > 4) {
		throw HipError("radix sort: at most 4 payload columns");
	}
	out = SortedRun {};
	out.npayload = npayload;
	const bool keys64 = pack.total_bits > 32;
	out.keys64 = keys64;
	const size_t kw = keys64 ? 8 : 4;
	if (n == 0) {
		out.keys = scratch.Alloc(kw);
		for (int p = 0; p < npayload; p++) {
			out.payload[p] = scratch.Alloc(8);
			out.payload_width[p] = payload_width[p];
		}
		return;
	}

	// Digit plan: the fewest passes of at most `radix_bits` bits covering the varying bits, split evenly.
	int B = std::max(pack.total_bits, 1);
	int rmax = tuning.radix_bits >= 11 ? 11 : 8;
	int P = tuning.passes_override > 0 ? tuning.passes_override : (B + rmax - 1) / rmax;
	if (P < 1) {
		P = 1;
	}
	if (P > 8) {
		P = 8;
	}
	PassList passes;
	passes.npasses = P;
	int shift = 0;
	int widest = 0;
	for (int p = 0; p < P; p++) {
		int bits = B / P + (p < B % P ? 1 : 0);
		passes.shift[p] = shift;
		passes.bits[p] = bits;
		shift += bits;
		widest = std::max(widest, bits);
	}
	const int R = widest <= 8 ? 8 : 11;
	if (widest > 11) {
		throw HipError("radix sort: digit wider than 11 bits requested");
	}
	out.passes = P;
	out.radix_bits = widest;
	const int RADIX = 1 << R;
	const uint64_t tile_rows = (uint64_t)EffectiveBlock(R, tuning) * EffectiveItems(R, tuning.items);
	const uint64_t tiles = (n + tile_rows - 1) / tile_rows;

	// Scratch: ping-pong keys and payloads; histogram; per-pass tile state + ticket.
	void *kbuf[2] = {scratch.Alloc(n * kw), P > 1 ? scratch.Alloc(n * kw) : nullptr};
	void *pbuf[kMaxPayload][2] = {};
	for (int p = 0; p < npayload; p++) {
		pbuf[p][0] = scratch.Alloc(n * payload_width[p]);
		pbuf[p][1] = P > 1 ? scratch.Alloc(n * payload_width[p]) : nullptr;
	}
	bool aligned8 = R == 8 && tuning.hist8 != nullptr && n_dev == nullptr;
	for (int p = 0; p < P; p++) {
		aligned8 = aligned8 && passes.bits[p] == 8 && passes.shift[p] == 8 * p;
	}
	uint32_t *hist = aligned8 ? const_cast<uint32_t *>(tuning.hist8)
	                          : static_cast<uint32_t *>(scratch.Alloc((size_t)P * RADIX * sizeof(uint32_t)));
	const uint64_t state_words = (uint64_t)P * tiles * RADIX + P;
	auto *state = static_cast<uint32_t *>(scratch.Alloc(state_words * sizeof(uint32_t)));
	ZeroSpans zero = tuning.zero_also;
	zero.Add(state, state_words);
	ZeroU32Spans(zero, stream);
	uint32_t *tickets = state + (uint64_t)P * tiles * RADIX;

	// Several key columns with real payload columns: pack once, beside the histograms, and start the passes
	// from the packed keys (pass 0 then reads one compact key; row-id payloads need the first-pass kernel).
	bool prepack = !aligned8 && R == 8 && pack.ncols > 1;
	for (int p = 0; p < npayload; p++) {
		prepack = prepack && payload_in[p] != nullptr;
	}
	if (tuning.narrow != 0) {
		// the narrowing load lives in the first pass that packs the key columns
		prepack = false;
		if (tuning.overflow == nullptr) {
			throw HipError("radix sort: payload narrowing needs an overflow word");
		}
		for (int p = 0; p < npayload; p++) {
			if (((tuning.narrow >> p) & 1u) != 0 && (payload_width[p] != 4 || payload_in[p] == nullptr)) {
				throw HipError("radix sort: a narrowed payload is a 4-byte carry of an INT64 column");
			}
		}
	}
	void *packed = nullptr;
	if (prepack) {
		packed = P > 1 ? kbuf[1] : scratch.Alloc(n * kw); // pass 0 writes kbuf[0]; pass 1 overwrites kbuf[1]
	}
	if (!aligned8) {
		FillU32(hist, 0u, (uint64_t)P * RADIX, stream);
		double bytes = prepack ? (double)n * kw : 0.0;
		for (int c = 0; c < pack.ncols; c++) {
			bytes += (double)n * pack.width[c];
		}
		ProfileScope prof("radix_hist", stream, bytes);
		uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 2048);
		if (prepack) {
			if (keys64) {
				hipLaunchKernelGGL(PackHistKernel<u64>, dim3((unsigned)blocks), dim3(256), 0, stream, pack, n, passes,
				                   (u64 *)packed, hist, n_dev);
			} else {
				hipLaunchKernelGGL(PackHistKernel<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, stream, pack, n,
				                   passes, (uint32_t *)packed, hist, n_dev);
			}
		} else if (R == 8) {
			hipLaunchKernelGGL(HistKernel<8>, dim3((unsigned)blocks), dim3(256), 0, stream, pack, n, passes, hist,
			                   n_dev);
		} else {
			hipLaunchKernelGGL(HistKernel<11>, dim3((unsigned)blocks), dim3(256), 0, stream, pack, n, passes, hist,
			                   n_dev);
		}
		VGPU_HIP(hipGetLastError());
	}

	const void *cur_keys = packed;
	const void *cur_pay[kMaxPayload];
	for (int p = 0; p < npayload; p++) {
		cur_pay[p] = payload_in[p];
	}
	// The caller may take over the last digit pass (an 8-bit pass that is not the first).
	const bool leave = tuning.leave_last && P >= 2 && R == 8 && n_dev == nullptr;
	const int run_passes = leave ? P - 1 : P;
	for (int pass = 0; pass < run_passes; pass++) {
		int slot = pass & 1;
		PayloadArgs pa;
		pa.n = npayload;
		for (int p = 0; p < npayload; p++) {
			pa.in[p] = cur_pay[p];
			pa.out[p] = pbuf[p][slot];
			pa.width[p] = payload_width[p];
		}
		if (pass == 0) {
			pa.narrow = tuning.narrow;
			pa.overflow = tuning.overflow;
		}
		uint32_t *st = state + (uint64_t)pass * tiles * RADIX;
		const uint32_t *gh = hist + (uint64_t)pass * RADIX;
		bool first = pass == 0 && !prepack;
		LaunchPass(R, keys64, first, pack, cur_keys, kbuf[slot], pa, n, passes.shift[pass], passes.bits[pass], gh,
		           st, tickets + pass, stream, tuning, n_dev);
		cur_keys = kbuf[slot];
		for (int p = 0; p < npayload; p++) {
			cur_pay[p] = pbuf[p][slot];
		}
	}
	out.keys = const_cast<void *>(cur_keys);
	for (int p = 0; p < npayload; p++) {
		out.payload[p] = const_cast<void *>(cur_pay[p]);
		out.payload_width[p] = payload_width[p];
	}
	out.max_packed = pack.max_packed != 0 ? pack.max_packed
	                                      : (pack.total_bits >= 64 ? ~0ull : (1ull << pack.total_bits) - 1);
	if (leave) {
		out.last_pending = true;
		out.last_shift = passes.shift[P - 1];
		out.last_bits = passes.bits[P - 1];
		out.last_hist = hist + (uint64_t)(P - 1) * RADIX;
	}
}

} // namespace vgpu
