// datagen.hip -- seeded, counter-based synthetic column generation on the device.
//
// Row i of a column is a pure function of (seed, stream id, i), so a 100M-row table is generated in one
// pass at HBM speed and any slice can be regenerated on the host bit-for-bit for the integer
// distributions (the oracle's generator uses the same mixer). Distributions: uniform integers in [lo, hi]
// (multiply-shift, bias <= range/2^64), uniform doubles in [0, 1) (top 53 bits), normal(0, 1)
// (Box-Muller), Zipf through an inverse-CDF table plus a rank -> key permutation.
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace vgpu {

__device__ __host__ __forceinline__ u64 MixCounter(u64 seed, u64 stream, u64 i) {
	u64 z = (seed ^ (stream * 0xD1B54A32D192ED03ull)) + (i + 1) * 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

namespace {

enum Dist : int { DIST_UNIFORM_INT = 0, DIST_UNIFORM_DOUBLE = 1, DIST_NORMAL = 2, DIST_ZIPF = 3, DIST_SEQUENCE = 4 };

__device__ __forceinline__ u64 MulHi(u64 a, u64 b) {
	return __umul64hi(a, b);
}

__device__ __forceinline__ void StoreAs(void *out, int vtype, uint64_t i, long long iv, double dv, bool is_float) {
	switch (vtype) {
	case VT_INT16:
		((short *)out)[i] = (short)iv;
		break;
	case VT_INT32:
		((int *)out)[i] = (int)iv;
		break;
	case VT_INT64:
		((long long *)out)[i] = iv;
		break;
	case VT_BOOLEAN:
		((unsigned char *)out)[i] = (unsigned char)(iv & 1);
		break;
	case VT_FLOAT32:
		((float *)out)[i] = is_float ? (float)dv : (float)iv;
		break;
	default:
		((double *)out)[i] = is_float ? dv : (double)iv;
		break;
	}
}

__global__ void GenKernel(void *out, int vtype, uint64_t n, int dist, long long lo, long long hi, u64 seed,
                          u64 stream_id, const double *cdf, uint64_t cdf_len, const long long *perm) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		u64 r = MixCounter(seed, stream_id, i);
		switch (dist) {
		case DIST_UNIFORM_INT: {
			u64 range = (u64)(hi - lo) + 1ull; // 0 means the full 2^64 range
			u64 off = range == 0 ? r : MulHi(r, range);
			StoreAs(out, vtype, i, (long long)((u64)lo + off), 0.0, false);
			break;
		}
		case DIST_UNIFORM_DOUBLE:
			StoreAs(out, vtype, i, 0, (double)(r >> 11) * 0x1.0p-53, true);
			break;
		case DIST_NORMAL: {
			u64 r2 = MixCounter(seed, stream_id ^ 0x5555555555555555ull, i);
			double u1 = ((double)(r >> 11) + 1.0) * 0x1.0p-53; // (0, 1]
			double u2 = (double)(r2 >> 11) * 0x1.0p-53;
			double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
			StoreAs(out, vtype, i, 0, z, true);
			break;
		}
		case DIST_ZIPF: {
			double u = (double)(r >> 11) * 0x1.0p-53;
			// first index with cdf[idx] > u
			uint64_t a = 0, b = cdf_len - 1;
			while (a < b) {
				uint64_t m = (a + b) >> 1;
				if (cdf[m] > u) {
					b = m;
				} else {
					a = m + 1;
				}
			}
			long long key = perm != nullptr ? perm[a] : (long long)a;
			StoreAs(out, vtype, i, key, 0.0, false);
			break;
		}
		default:
			StoreAs(out, vtype, i, (long long)i + lo, 0.0, false);
			break;
		}
	}
}

} // namespace

void GenerateColumn(void *out, int vtype, uint64_t n, int dist, long long lo, long long hi, uint64_t seed,
                    uint64_t stream_id, const double *cdf, uint64_t cdf_len, const long long *perm,
                    hipStream_t stream) {
	if (n == 0) {
		return;
	}
	uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 16384);
	hipLaunchKernelGGL(GenKernel, dim3((unsigned)blocks), dim3(256), 0, stream, out, vtype, n, dist, lo, hi,
	                   (u64)seed, (u64)stream_id, cdf, cdf_len, perm);
	VGPU_HIP(hipGetLastError());
}

} // namespace vgpu
