// scan.hip -- running (cumulative) sum of one column with a carried state, for the streaming pipeline's
// RUNNING_SUM window primitive (cuda_engine/src/gpu_stream_pipeline.cu:133-167): out[i] = state + x[0] +
// ... + x[i], then state = out[n - 1], so consecutive batches continue one sum.
//
// Three short kernels on the caller's stream, deterministic for every type: per-tile totals, one block
// scanning the tile totals on top of the carried state, then each tile's inclusive scan plus its offset.
// Integer sums wrap in two's complement (computed unsigned); floating sums add in a fixed tree order.
#include "common.h"
#include "kernels.h"

namespace vgpu {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

template <typename T>
struct AccOf {
	using type = T;
};
template <>
struct AccOf<int16_t> {
	using type = uint32_t;
};
template <>
struct AccOf<int32_t> {
	using type = uint32_t;
};
template <>
struct AccOf<int64_t> {
	using type = u64;
};

//! Inclusive and exclusive scans over the block of one value per thread, with no subtraction (exact order
//! for floating types); *total receives the block sum.
template <typename A>
__device__ __forceinline__ void BlockScan(A v, A *lds, A *inclusive, A *exclusive, A *total) {
	const int lane = LaneId();
	const int wave = threadIdx.x / kWave;
	for (int o = 1; o < kWave; o <<= 1) {
		A up = __shfl_up(v, o);
		if (lane >= o) {
			v = v + up;
		}
	}
	if (lane == kWave - 1) {
		lds[wave] = v;
	}
	__syncthreads();
	A before = A(0);
	A all = A(0);
	for (int w = 0; w < kScanBlock / kWave; w++) {
		if (w < wave) {
			before = before + lds[w];
		}
		all = all + lds[w];
	}
	__syncthreads();
	const A incl = before + v;
	const A up = __shfl_up(incl, 1);
	*inclusive = incl;
	*exclusive = lane == 0 ? before : up;
	*total = all;
}

template <typename T>
__global__ void __launch_bounds__(kScanBlock) TileTotalsKernel(const T *x, uint64_t n, typename AccOf<T>::type *tiles) {
	using A = typename AccOf<T>::type;
	__shared__ A lds[kScanBlock / kWave];
	const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
	A s = A(0);
#pragma unroll
	for (int j = 0; j < kScanItems; j++) {
		if (base + j < n) {
			s = s + (A)x[base + j];
		}
	}
	A incl, excl, total;
	BlockScan(s, lds, &incl, &excl, &total);
	if (threadIdx.x == 0) {
		tiles[blockIdx.x] = total;
	}
}

//! One block: tiles[t] becomes state + (sum of the tiles before t), in tile order.
template <typename T>
__global__ void __launch_bounds__(kScanBlock) TileOffsetsKernel(typename AccOf<T>::type *tiles, uint64_t ntiles,
                                                                const T *state) {
	using A = typename AccOf<T>::type;
	__shared__ A lds[kScanBlock / kWave];
	A carry = (A)*state;
	for (uint64_t first = 0; first < ntiles; first += kScanBlock) {
		const uint64_t t = first + threadIdx.x;
		const A v = t < ntiles ? tiles[t] : A(0);
		A incl, excl, total;
		BlockScan(v, lds, &incl, &excl, &total);
		if (t < ntiles) {
			tiles[t] = carry + excl;
		}
		carry = carry + total;
		__syncthreads();
	}
}

template <typename T>
__global__ void __launch_bounds__(kScanBlock) TileScanKernel(T *x, uint64_t n, const typename AccOf<T>::type *offsets,
                                                             T *state) {
	using A = typename AccOf<T>::type;
	__shared__ A lds[kScanBlock / kWave];
	const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
	A p[kScanItems];
	A run = A(0);
#pragma unroll
	for (int j = 0; j < kScanItems; j++) {
		run = run + (base + j < n ? (A)x[base + j] : A(0));
		p[j] = run;
	}
	A incl, excl, total;
	BlockScan(run, lds, &incl, &excl, &total);
	const A offset = offsets[blockIdx.x];
#pragma unroll
	for (int j = 0; j < kScanItems; j++) {
		if (base + j < n) {
			const A v = offset + (excl + p[j]);
			x[base + j] = (T)v;
			if (base + j == n - 1) {
				*state = (T)v;
			}
		}
	}
}

} // namespace

size_t RunningSumScratchBytes(uint64_t n) {
	return ((n + kScanTile - 1) / kScanTile + 1) * sizeof(u64);
}

template <typename T>
static void RunningSumT(T *data, uint64_t n, T *state, void *scratch, hipStream_t stream) {
	using A = typename AccOf<T>::type;
	const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
	auto *acc = static_cast<A *>(scratch);
	hipLaunchKernelGGL((TileTotalsKernel<T>), dim3((unsigned)tiles), dim3(kScanBlock), 0, stream, data, n, acc);
	hipLaunchKernelGGL((TileOffsetsKernel<T>), dim3(1), dim3(kScanBlock), 0, stream, acc, tiles, state);
	hipLaunchKernelGGL((TileScanKernel<T>), dim3((unsigned)tiles), dim3(kScanBlock), 0, stream, data, n, acc, state);
	VGPU_HIP(hipGetLastError());
}

void RunningSum(int vtype, void *data, uint64_t n, void *state, void *scratch, size_t scratch_bytes,
                hipStream_t stream) {
	if (n == 0) {
		return;
	}
	if (scratch_bytes < RunningSumScratchBytes(n)) {
		throw HipError("RunningSum: scratch too small");
	}
	switch (vtype) {
	case VT_INT16:
		RunningSumT(static_cast<int16_t *>(data), n, static_cast<int16_t *>(state), scratch, stream);
		break;
	case VT_INT32:
		RunningSumT(static_cast<int32_t *>(data), n, static_cast<int32_t *>(state), scratch, stream);
		break;
	case VT_INT64:
		RunningSumT(static_cast<int64_t *>(data), n, static_cast<int64_t *>(state), scratch, stream);
		break;
	case VT_FLOAT32:
		RunningSumT(static_cast<float *>(data), n, static_cast<float *>(state), scratch, stream);
		break;
	case VT_FLOAT64:
		RunningSumT(static_cast<double *>(data), n, static_cast<double *>(state), scratch, stream);
		break;
	default:
		throw HipError("RunningSum: unsupported type for RUNNING_SUM");
	}
}

} // namespace vgpu
