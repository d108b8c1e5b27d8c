// extended.cpp -- device helpers of the opt-in semantics (see extended.h). Every kernel is a grid-stride
// elementwise pass: each row is read once and written once, so all of them are HBM-bound copies.
#include "extended.h"

#include "../hip/onesweep.h"

#include <algorithm>
#include <stdexcept>

namespace vgpu {

namespace {

unsigned Grid(uint64_t n) {
	const uint64_t g = (n + 255) / 256;
	return (unsigned)std::min<uint64_t>(std::max<uint64_t>(g, 1), 8192);
}

#define VGPU_GRID_STRIDE(i, n) \
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// Canonical value under DuckDB's equality, then the sign-magnitude -> two's-complement order flip.
template <typename U, int kBits>
__device__ __forceinline__ U OrderedFloatBits(U b) {
	constexpr U kSign = (U)1 << (kBits - 1);
	constexpr U kExp = kBits == 64 ? (U)0x7FF0000000000000ull : (U)0x7F800000u;
	constexpr U kQuietNaN = kBits == 64 ? (U)0x7FF8000000000000ull : (U)0x7FC00000u;
	const U mag = b & ~kSign;
	if (mag == 0) {
		b = 0; // -0.0 == +0.0
	} else if ((mag & kExp) == kExp && (mag & ~kExp) != 0) {
		b = kQuietNaN; // every NaN is one value, above +inf
	}
	return (b & kSign) ? (b ^ (kSign - 1)) : b;
}

template <typename U, int kBits>
__global__ void FloatToOrderedKernel(const U *src, uint64_t n, U *dst) {
	VGPU_GRID_STRIDE(i, n) {
		dst[i] = OrderedFloatBits<U, kBits>(src[i]);
	}
}

template <typename U, int kBits>
__global__ void OrderedToFloatKernel(const U *src, uint64_t n, U *dst) {
	constexpr U kSign = (U)1 << (kBits - 1);
	VGPU_GRID_STRIDE(i, n) {
		const U s = src[i];
		dst[i] = (s & kSign) ? (s ^ (kSign - 1)) : s;
	}
}

template <typename T>
__global__ void MaskNullsKernel(const T *src, const uint8_t *valid, T fill, uint64_t n, T *dst) {
	VGPU_GRID_STRIDE(i, n) {
		dst[i] = valid[i] != 0 ? src[i] : fill;
	}
}

__global__ void NullFlagsKernel(const uint8_t *valid, uint64_t n, uint8_t *dst) {
	VGPU_GRID_STRIDE(i, n) {
		dst[i] = valid[i] != 0 ? 0 : 1;
	}
}

__global__ void ValidAsInt32Kernel(const uint8_t *valid, uint64_t n, int32_t *dst) {
	VGPU_GRID_STRIDE(i, n) {
		dst[i] = valid[i] != 0 ? 1 : 0;
	}
}

__global__ void ValidFromFlagsKernel(const uint8_t *flag, uint64_t n, uint8_t *valid) {
	VGPU_GRID_STRIDE(i, n) {
		valid[i] = flag[i] == 0 ? 1 : 0;
	}
}

template <typename T>
__global__ void ValidUnlessEqualKernel(const T *values, T sentinel, uint64_t n, uint8_t *valid) {
	VGPU_GRID_STRIDE(i, n) {
		valid[i] = values[i] != sentinel ? 1 : 0;
	}
}

//! The 128-bit two's-complement value {lo, hi} rounded to double: |v| as hi * 2^64 + lo, sign applied.
__device__ __forceinline__ double HugeToDouble(u64 lo, u64 hi) {
	const bool neg = (int64_t)hi < 0;
	if (neg) {
		lo = ~lo + 1;
		hi = ~hi + (lo == 0 ? 1 : 0);
	}
	const double v = (double)hi * 18446744073709551616.0 + (double)lo;
	return neg ? -v : v;
}

__global__ void FinishNullableKernel(int kind, int width, void *out, uint8_t *out_valid, const void *sum,
                                     int sum_is_huge, const int64_t *count, uint64_t groups) {
	VGPU_GRID_STRIDE(g, groups) {
		const int64_t c = count[g];
		out_valid[g] = c > 0 ? 1 : 0;
		if (kind == AGG_AVG) {
			double v = 0.0;
			if (c > 0) {
				const double s = sum_is_huge ? HugeToDouble(((const u64 *)sum)[2 * g], ((const u64 *)sum)[2 * g + 1])
				                             : ((const double *)sum)[g];
				v = s / (double)c;
			}
			((double *)out)[g] = v;
		} else if (c == 0) {
			uint8_t *p = (uint8_t *)out + g * (uint64_t)width;
			for (int b = 0; b < width; b++) {
				p[b] = 0;
			}
		}
	}
}

__global__ void PackKeysKernel(KeyPack pack, uint64_t n, u64 *dst) {
	VGPU_GRID_STRIDE(i, n) {
		dst[i] = sortpass::PackRow(pack, i);
	}
}

} // namespace

void FloatKeyToOrdered(const void *src, int vtype, uint64_t n, void *dst, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	if (vtype == VT_FLOAT64) {
		hipLaunchKernelGGL((FloatToOrderedKernel<u64, 64>), dim3(Grid(n)), dim3(256), 0, stream, (const u64 *)src, n,
		                   (u64 *)dst);
	} else if (vtype == VT_FLOAT32) {
		hipLaunchKernelGGL((FloatToOrderedKernel<uint32_t, 32>), dim3(Grid(n)), dim3(256), 0, stream,
		                   (const uint32_t *)src, n, (uint32_t *)dst);
	} else {
		throw std::runtime_error("float key transform: FLOAT or DOUBLE column expected");
	}
	VGPU_HIP(hipGetLastError());
}

void OrderedToFloatKey(const void *src, int vtype, uint64_t n, void *dst, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	if (vtype == VT_FLOAT64) {
		hipLaunchKernelGGL((OrderedToFloatKernel<u64, 64>), dim3(Grid(n)), dim3(256), 0, stream, (const u64 *)src, n,
		                   (u64 *)dst);
	} else if (vtype == VT_FLOAT32) {
		hipLaunchKernelGGL((OrderedToFloatKernel<uint32_t, 32>), dim3(Grid(n)), dim3(256), 0, stream,
		                   (const uint32_t *)src, n, (uint32_t *)dst);
	} else {
		throw std::runtime_error("float key transform: FLOAT or DOUBLE column expected");
	}
	VGPU_HIP(hipGetLastError());
}

void MaskNulls(const void *src, const uint8_t *valid, int width, uint64_t fill, uint64_t n, void *dst,
               hipStream_t stream) {
	if (n == 0) {
		return;
	}
	switch (width) {
	case 1:
		hipLaunchKernelGGL(MaskNullsKernel<uint8_t>, dim3(Grid(n)), dim3(256), 0, stream, (const uint8_t *)src, valid,
		                   (uint8_t)fill, n, (uint8_t *)dst);
		break;
	case 2:
		hipLaunchKernelGGL(MaskNullsKernel<uint16_t>, dim3(Grid(n)), dim3(256), 0, stream, (const uint16_t *)src,
		                   valid, (uint16_t)fill, n, (uint16_t *)dst);
		break;
	case 4:
		hipLaunchKernelGGL(MaskNullsKernel<uint32_t>, dim3(Grid(n)), dim3(256), 0, stream, (const uint32_t *)src,
		                   valid, (uint32_t)fill, n, (uint32_t *)dst);
		break;
	case 8:
		hipLaunchKernelGGL(MaskNullsKernel<u64>, dim3(Grid(n)), dim3(256), 0, stream, (const u64 *)src, valid,
		                   (u64)fill, n, (u64 *)dst);
		break;
	default:
		throw std::runtime_error("null mask: 1, 2, 4 or 8-byte values");
	}
	VGPU_HIP(hipGetLastError());
}

void NullFlags(const uint8_t *valid, uint64_t n, uint8_t *dst, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(NullFlagsKernel, dim3(Grid(n)), dim3(256), 0, stream, valid, n, dst);
	VGPU_HIP(hipGetLastError());
}

void ValidAsInt32(const uint8_t *valid, uint64_t n, int32_t *dst, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(ValidAsInt32Kernel, dim3(Grid(n)), dim3(256), 0, stream, valid, n, dst);
	VGPU_HIP(hipGetLastError());
}

void ValidFromNullFlags(const uint8_t *flag, uint64_t n, uint8_t *valid, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(ValidFromFlagsKernel, dim3(Grid(n)), dim3(256), 0, stream, flag, n, valid);
	VGPU_HIP(hipGetLastError());
}

void ValidUnlessEqual(const void *values, int width, uint64_t sentinel, uint64_t n, uint8_t *valid,
                      hipStream_t stream) {
	if (n == 0) {
		return;
	}
	if (width == 8) {
		hipLaunchKernelGGL(ValidUnlessEqualKernel<u64>, dim3(Grid(n)), dim3(256), 0, stream, (const u64 *)values,
		                   (u64)sentinel, n, valid);
	} else if (width == 4) {
		hipLaunchKernelGGL(ValidUnlessEqualKernel<uint32_t>, dim3(Grid(n)), dim3(256), 0, stream,
		                   (const uint32_t *)values, (uint32_t)sentinel, n, valid);
	} else {
		throw std::runtime_error("sentinel validity: 4 or 8-byte values");
	}
	VGPU_HIP(hipGetLastError());
}

void FinishNullableAggregate(int kind, int out_type, void *out, uint8_t *out_valid, const void *sum, int sum_type,
                             const int64_t *count, uint64_t groups, hipStream_t stream) {
	if (groups == 0) {
		return;
	}
	if (kind == AGG_AVG && (sum == nullptr || out_type != VT_FLOAT64)) {
		throw std::runtime_error("nullable AVG: needs its SUM and a DOUBLE output");
	}
	hipLaunchKernelGGL(FinishNullableKernel, dim3(Grid(groups)), dim3(256), 0, stream, kind, TypeWidth(out_type), out,
	                   out_valid, sum, sum_type == VT_HUGEINT ? 1 : 0, count, groups);
	VGPU_HIP(hipGetLastError());
}

void PackKeys(const KeyPack &pack, uint64_t n, uint64_t *dst, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(PackKeysKernel, dim3(Grid(n)), dim3(256), 0, stream, pack, n, (u64 *)dst);
	VGPU_HIP(hipGetLastError());
}

} // namespace vgpu
