// misc.hip -- small device primitives: fills, gathers, validity unpack, cross-product expansion, the
// order-preserving int32 threshold filter, type helpers and per-launch HIP-event profiling.
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace vgpu {

int TypeWidth(int vtype) {
	switch (vtype) {
	case VT_INT16:
		return 2;
	case VT_INT32:
		return 4;
	case VT_INT64:
		return 8;
	case VT_FLOAT32:
		return 4;
	case VT_FLOAT64:
		return 8;
	case VT_BOOLEAN:
		return 1;
	case VT_HUGEINT:
		return 16;
	default:
		return -1;
	}
}

bool TypeIsSigned(int vtype) {
	return vtype == VT_INT16 || vtype == VT_INT32 || vtype == VT_INT64 || vtype == VT_HUGEINT;
}

bool TypeIsInteger(int vtype) {
	return vtype == VT_INT16 || vtype == VT_INT32 || vtype == VT_INT64 || vtype == VT_BOOLEAN;
}

void Check(hipError_t status, const char *what) {
	if (status != hipSuccess) {
		throw HipError(std::string(what) + " failed: " + hipGetErrorString(status));
	}
}

// ------------------------------------------------------------------ profiling

namespace {
struct ProfRecord {
	std::string name;
	hipEvent_t start;
	hipEvent_t stop;
	double bytes;
};
std::mutex g_prof_mutex;
bool g_prof_on = false;
std::string g_prof_only; // when set, only scopes of this name are timed
std::vector<ProfRecord> g_prof_pending;
std::map<std::string, ProfileEntry> g_prof_totals;
//! Events of read-out records, reused by later scopes (creating two events per launch cost microseconds of
//! host time inside the timed region).
std::vector<hipEvent_t> g_prof_free;

hipEvent_t TakeProfEvent() {
	{
		std::lock_guard<std::mutex> lock(g_prof_mutex);
		if (!g_prof_free.empty()) {
			hipEvent_t e = g_prof_free.back();
			g_prof_free.pop_back();
			return e;
		}
	}
	hipEvent_t e = nullptr;
	VGPU_HIP(hipEventCreate(&e));
	return e;
}
} // namespace

void ProfileEnable(bool on) {
	std::lock_guard<std::mutex> lock(g_prof_mutex);
	g_prof_on = on;
}

bool ProfileEnabled() {
	std::lock_guard<std::mutex> lock(g_prof_mutex);
	return g_prof_on;
}

void ProfileOnly(const char *name) {
	std::lock_guard<std::mutex> lock(g_prof_mutex);
	g_prof_only = name != nullptr ? name : "";
}

ProfileScope::ProfileScope(const char *name, hipStream_t stream, double bytes)
    : name_(name), stream_(stream), bytes_(bytes) {
	{
		std::lock_guard<std::mutex> lock(g_prof_mutex);
		if (!g_prof_on || (!g_prof_only.empty() && g_prof_only != name_)) {
			return;
		}
	}
	// launches captured into a graph are not timed here (the events would never be recorded)
	hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
	if (hipStreamIsCapturing(stream_, &capture) != hipSuccess || capture != hipStreamCaptureStatusNone) {
		return;
	}
	start_ = TakeProfEvent();
	stop_ = TakeProfEvent();
	VGPU_HIP(hipEventRecord(start_, stream_));
}

ProfileScope::~ProfileScope() {
	if (start_ == nullptr) {
		return;
	}
	if (hipEventRecord(stop_, stream_) != hipSuccess) {
		(void)0;
		(void)hipEventDestroy(start_);
		(void)hipEventDestroy(stop_);
		return;
	}
	std::lock_guard<std::mutex> lock(g_prof_mutex);
	g_prof_pending.push_back({name_, start_, stop_, bytes_});
}

std::vector<ProfileEntry> ProfileSnapshot() {
	std::lock_guard<std::mutex> lock(g_prof_mutex);
	for (auto &rec : g_prof_pending) {
		(void)hipEventSynchronize(rec.stop);
		float ms = 0;
		(void)hipEventElapsedTime(&ms, rec.start, rec.stop);
		auto &entry = g_prof_totals[rec.name];
		entry.name = rec.name;
		entry.total_ms += ms;
		entry.launches += 1;
		entry.bytes += rec.bytes;
		g_prof_free.push_back(rec.start);
		g_prof_free.push_back(rec.stop);
	}
	g_prof_pending.clear();
	std::vector<ProfileEntry> out;
	for (auto &kv : g_prof_totals) {
		out.push_back(kv.second);
	}
	return out;
}

void ProfileReset() {
	ProfileSnapshot();
	std::lock_guard<std::mutex> lock(g_prof_mutex);
	g_prof_totals.clear();
}

// ------------------------------------------------------------------ read-back

void ReadBack(void *host, const void *dev, size_t bytes, hipStream_t stream) {
	constexpr size_t kStage = 4096;
	thread_local void *stage = nullptr;
	if (bytes == 0) {
		return;
	}
	if (stage == nullptr && bytes <= kStage) {
		if (hipHostMalloc(&stage, kStage, hipHostMallocDefault) != hipSuccess) {
			(void)hipGetLastError();
			stage = nullptr;
		}
	}
	void *dst = (bytes <= kStage && stage != nullptr) ? stage : host;
	VGPU_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, stream));
	while (true) {
		const hipError_t st = hipStreamQuery(stream);
		if (st == hipSuccess) {
			break;
		}
		if (st != hipErrorNotReady) {
			VGPU_HIP(st);
		}
		(void)hipGetLastError();
	}
	if (dst != host) {
		std::memcpy(host, dst, bytes);
	}
}

// ------------------------------------------------------------------ fills

namespace {
unsigned GridFor(uint64_t n, unsigned block = 256, unsigned cap = 8192) {
	uint64_t g = (n + block - 1) / block;
	if (g == 0) {
		g = 1;
	}
	return (unsigned)std::min<uint64_t>(g, cap);
}
} // namespace

void FillU32(uint32_t *dst, uint32_t value, uint64_t n, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(FillKernel<uint32_t>, dim3(GridFor(n)), dim3(256), 0, stream, dst, value, n);
	VGPU_HIP(hipGetLastError());
}

void FillU64(uint64_t *dst, uint64_t value, uint64_t n, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(FillKernel<u64>, dim3(GridFor(n)), dim3(256), 0, stream, (u64 *)dst, (u64)value, n);
	VGPU_HIP(hipGetLastError());
}

void ZeroSpans::Add(uint32_t *p, uint64_t n) {
	if (p == nullptr || n == 0) {
		return;
	}
	if (count == kMax) {
		throw HipError("ZeroSpans: more than four spans");
	}
	ptr[count] = p;
	words[count] = n;
	count++;
}

namespace {
__global__ void __launch_bounds__(256) ZeroSpansKernel(ZeroSpans spans) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	for (int k = 0; k < spans.count; k++) {
		uint32_t *p = spans.ptr[k];
		const uint64_t n = spans.words[k];
		for (uint64_t i = first; i < n; i += stride) {
			p[i] = 0;
		}
	}
}
} // namespace

void ZeroU32Spans(const ZeroSpans &spans, hipStream_t stream) {
	uint64_t most = 0;
	for (int k = 0; k < spans.count; k++) {
		most = std::max<uint64_t>(most, spans.words[k]);
	}
	if (most == 0) {
		return;
	}
	hipLaunchKernelGGL(ZeroSpansKernel, dim3(GridFor(most)), dim3(256), 0, stream, spans);
	VGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ device copy

namespace {
typedef unsigned int CopyVec __attribute__((ext_vector_type(4)));
//! 16-byte words per thread of the device copy: each block moves one contiguous 16 KiB chunk (four
//! 16-byte loads in flight per lane, then four stores). Many small contiguous chunks keep every channel
//! busy; a grid-stride loop over the same bytes runs measurably slower on this part.
constexpr int kCopyItems = 4;

__global__ void __launch_bounds__(256) CopyChunkKernel(const CopyVec *src, CopyVec *dst, uint64_t n16,
                                                       const uint8_t *tail_src, uint8_t *tail_dst,
                                                       uint32_t tail) {
	const uint64_t base = (uint64_t)blockIdx.x * (256 * kCopyItems) + threadIdx.x;
	CopyVec v[kCopyItems];
#pragma unroll
	for (int u = 0; u < kCopyItems; u++) {
		const uint64_t i = base + (uint64_t)u * 256;
		v[u] = i < n16 ? src[i] : CopyVec {0u, 0u, 0u, 0u};
	}
#pragma unroll
	for (int u = 0; u < kCopyItems; u++) {
		const uint64_t i = base + (uint64_t)u * 256;
		if (i < n16) {
			dst[i] = v[u];
		}
	}
	if (blockIdx.x == 0 && threadIdx.x < tail) {
		tail_dst[threadIdx.x] = tail_src[threadIdx.x];
	}
}
} // namespace

void CopyDevice(void *dst, const void *src, uint64_t bytes, hipStream_t stream) {
	if (bytes == 0) {
		return;
	}
	if ((((uintptr_t)dst | (uintptr_t)src) & 15) != 0) {
		VGPU_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream));
		return;
	}
	const uint64_t n16 = bytes / 16;
	const uint32_t tail = (uint32_t)(bytes % 16);
	const uint64_t blocks = std::max<uint64_t>((n16 + 256 * kCopyItems - 1) / (256 * kCopyItems), 1);
	if (blocks > 0x7fffffffull) {
		throw HipError("CopyDevice: copy too large for one launch");
	}
	hipLaunchKernelGGL(CopyChunkKernel, dim3((unsigned)blocks), dim3(256), 0, stream,
	                   static_cast<const CopyVec *>(src), static_cast<CopyVec *>(dst), n16,
	                   static_cast<const uint8_t *>(src) + n16 * 16, static_cast<uint8_t *>(dst) + n16 * 16, tail);
	VGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ gather

namespace {
struct alignas(16) U128 {
	u64 lo;
	u64 hi;
};

template <typename T, typename I>
__global__ void GatherKernel(const T *src, const I *idx, uint64_t n, T *out) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		out[i] = src[idx[i]];
	}
}

template <typename T>
void GatherT(const void *src, const void *idx, bool idx64, uint64_t n, void *out, hipStream_t stream) {
	if (idx64) {
		hipLaunchKernelGGL((GatherKernel<T, u64>), dim3(GridFor(n)), dim3(256), 0, stream, (const T *)src,
		                   (const u64 *)idx, n, (T *)out);
	} else {
		hipLaunchKernelGGL((GatherKernel<T, uint32_t>), dim3(GridFor(n)), dim3(256), 0, stream, (const T *)src,
		                   (const uint32_t *)idx, n, (T *)out);
	}
	VGPU_HIP(hipGetLastError());
}
} // namespace

void Gather(const void *src, int width, const void *idx, bool idx64, uint64_t n, void *out, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	ProfileScope prof("gather", stream, (double)n * (2.0 * width + (idx64 ? 8 : 4)));
	switch (width) {
	case 1:
		GatherT<uint8_t>(src, idx, idx64, n, out, stream);
		break;
	case 2:
		GatherT<uint16_t>(src, idx, idx64, n, out, stream);
		break;
	case 4:
		GatherT<uint32_t>(src, idx, idx64, n, out, stream);
		break;
	case 8:
		GatherT<u64>(src, idx, idx64, n, out, stream);
		break;
	case 16:
		GatherT<U128>(src, idx, idx64, n, out, stream);
		break;
	default:
		throw HipError("Gather: unsupported element width");
	}
}

// ------------------------------------------------------------------ validity

namespace {
__global__ void UnpackValidityKernel(const uint8_t *packed, uint8_t *out, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		out[i] = (packed[i >> 3] >> (i & 7)) & 1u;
	}
}
} // namespace

void UnpackValidity(const uint8_t *packed, uint8_t *out, uint64_t n, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(UnpackValidityKernel, dim3(GridFor(n)), dim3(256), 0, stream, packed, out, n);
	VGPU_HIP(hipGetLastError());
}

namespace {
//! One thread per row: a wave's 64 validity bytes become one 64-bit ballot, i.e. 8 packed bytes, written
//! by the lanes that own them (lane 8k writes byte k), so the packed output is written once, coalesced.
__global__ void PackValidityKernel(const uint8_t *dense, uint8_t *packed, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	const int lane = threadIdx.x & 63;
	for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
		const uint64_t i = base + threadIdx.x;
		const bool v = i < n && dense[i] != 0;
		const u64 bits = __ballot(v);
		if ((lane & 7) == 0 && i < n) {
			packed[i >> 3] = (uint8_t)(bits >> lane);
		}
	}
}

__global__ void FillU8Kernel(uint8_t *dst, uint8_t value, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		dst[i] = value;
	}
}
} // namespace

void PackValidity(const uint8_t *dense, uint8_t *packed, uint64_t n, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	// Block size is a multiple of 64 and the grid stride is a multiple of the block, so every wave's
	// 64 rows start on a byte boundary of the packed output (row index multiple of 64).
	hipLaunchKernelGGL(PackValidityKernel, dim3(GridFor(n)), dim3(256), 0, stream, dense, packed, n);
	VGPU_HIP(hipGetLastError());
}

void FillU8(uint8_t *dst, uint8_t value, uint64_t n, hipStream_t stream) {
	if (n == 0) {
		return;
	}
	hipLaunchKernelGGL(FillU8Kernel, dim3(GridFor(n)), dim3(256), 0, stream, dst, value, n);
	VGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ cross product expansion

namespace {
//! Left-major pairing: output row i is (left i / m, right i % m). A left column broadcasts one value to m
//! consecutive outputs; a right column repeats the whole (cache-resident) right side.
template <typename T>
__global__ void CrossExpandKernel(const T *src, uint64_t m, bool from_left, uint64_t total, T *dst) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
		dst[i] = src[from_left ? i / m : i % m];
	}
}
template <typename T>
void CrossT(const void *src, uint64_t m, bool from_left, uint64_t total, void *dst, hipStream_t stream) {
	hipLaunchKernelGGL(CrossExpandKernel<T>, dim3(GridFor(total, 256, 16384)), dim3(256), 0, stream,
	                   (const T *)src, m, from_left, total, (T *)dst);
	VGPU_HIP(hipGetLastError());
}
} // namespace

void CrossExpand(const void *src, int width, uint64_t m, bool from_left, uint64_t total, void *dst,
                 hipStream_t stream) {
	if (total == 0) {
		return;
	}
	ProfileScope prof("cross_expand", stream, (double)total * width);
	switch (width) {
	case 1:
		CrossT<uint8_t>(src, m, from_left, total, dst, stream);
		break;
	case 2:
		CrossT<uint16_t>(src, m, from_left, total, dst, stream);
		break;
	case 4:
		CrossT<uint32_t>(src, m, from_left, total, dst, stream);
		break;
	case 8:
		CrossT<u64>(src, m, from_left, total, dst, stream);
		break;
	case 16:
		CrossT<U128>(src, m, from_left, total, dst, stream);
		break;
	default:
		throw HipError("CrossExpand: unsupported element width");
	}
}

// ------------------------------------------------------------------ FilterGreaterThanInt32

namespace {
constexpr int kFgBlock = 256;
} // namespace

namespace {
//! Order-preserving compaction of the keys with lo <= x <= hi (signed). Tile = 256 threads x kItems rows,
//! wave-striped (item k of a thread is row base + k * 256 + tid, so every load is one coalesced line per
//! wave); a row's slot in the tile is the kept count of the (item, wave) groups before it (ballots, one
//! small scan in LDS) plus its rank among its wave's kept lanes, so each wave writes contiguous runs; the
//! tile's global offset comes from a decoupled look-back.
//! 128 bytes per thread: tiles of 4096 (8-byte keys) or 8192 (4-byte) rows.
template <typename T>
constexpr int KeepRangeItems() {
	return (int)(128 / sizeof(T));
}

template <typename T, int BLOCK>
__global__ void __launch_bounds__(BLOCK) KeepRangeKernel(const T *in, uint64_t n, long long lo, long long hi,
                                                        T *out, u64 *state, uint32_t *ticket, u64 *count_out,
                                                        uint64_t ntiles) {
	constexpr int kItems = KeepRangeItems<T>();
	constexpr int kWaves = BLOCK / kWave;
	__shared__ uint32_t s_tile;
	__shared__ uint32_t s_off[kItems * kWaves];
	__shared__ u64 s_scan[kWaves + 1];
	__shared__ u64 s_prefix;
	const int tid = threadIdx.x;
	const int lane = tid & (kWave - 1);
	const int wave = tid / kWave;
	if (tid == 0) {
		s_tile = FetchAddAgent(ticket, 1u);
	}
	__syncthreads();
	const uint32_t tile = s_tile;
	const uint64_t base = (uint64_t)tile * BLOCK * kItems;
	T v[kItems];
	u64 ball[kItems];
#pragma unroll
	for (int k = 0; k < kItems; k++) {
		const uint64_t row = base + (uint64_t)k * BLOCK + tid;
		v[k] = row < n ? in[row] : (T)0;
	}
#pragma unroll
	for (int k = 0; k < kItems; k++) {
		const uint64_t row = base + (uint64_t)k * BLOCK + tid;
		const long long x = (long long)v[k];
		ball[k] = __ballot(row < n && x >= lo && x <= hi);
		if (lane == 0) {
			s_off[k * kWaves + wave] = (uint32_t)Popc64(ball[k]);
		}
	}
	__syncthreads();
	static_assert(kItems * kWaves <= BLOCK, "one (item, wave) count per thread");
	{
		// exclusive scan of the (item, wave) counts, one per thread
		const u64 mine = tid < kItems * kWaves ? s_off[tid] : 0ull;
		u64 run = 0;
		const u64 ex = BlockExclusiveScan<u64, BLOCK>(mine, s_scan, &run);
		if (tid < kItems * kWaves) {
			s_off[tid] = (uint32_t)ex;
		}
		if (tid == 0) {
			StoreAgent(state + tile, LbPack(tile == 0 ? kLbInclusive : kLbAggregate, run));
			s_prefix = run; // tile total until the look-back below replaces it
		}
	}
	__syncthreads();
	if (wave == 0) {
		const u64 total = s_prefix;
		const u64 prefix = tile == 0 ? 0ull : LookbackExclusiveWave(state, tile);
		if (lane == 0) {
			if (tile > 0) {
				StoreAgent(state + tile, LbPack(kLbInclusive, prefix + total));
			}
			if (tile == ntiles - 1) {
				*count_out = prefix + total;
			}
		}
		__builtin_amdgcn_wave_barrier();
		if (lane == 0) {
			s_prefix = prefix;
		}
	}
	__syncthreads();
	const u64 tile_out = s_prefix;
	const u64 lt = LaneMaskLt();
#pragma unroll
	for (int k = 0; k < kItems; k++) {
		if ((ball[k] >> lane) & 1ull) {
			out[tile_out + s_off[k * kWaves + wave] + (u64)Popc64(ball[k] & lt)] = v[k];
		}
	}
}

//! The same compaction for a 16-byte aligned input: each lane loads 16-byte vectors of kVec consecutive rows (a
//! wave reads one contiguous 1 KiB line per load), a kept row's rank in its wave comes from one ballot per
//! vector component (mbcnt of the ballots below the lane plus the lane's own earlier components), kept rows are
//! staged in LDS in tile order and written out as one contiguous run per tile (every wave store a full
//! 256-byte line). Same tile, look-back words and ticket as KeepRangeKernel. Measured on 2^28 INT32 rows, half
//! kept: 0.43 -> 0.41 ms; a persistent form that loads its next tile during the current one's look-back and
//! write-out ran 0.60 ms (the ticket and the next tile's loads then sit on each block's serial chain).
template <typename T, int BLOCK>
__global__ void __launch_bounds__(BLOCK) KeepRangeVecKernel(const T *in, uint64_t n, long long lo, long long hi,
                                                           T *out, u64 *state, uint32_t *ticket, u64 *count_out,
                                                           uint64_t ntiles) {
	constexpr int kItems = KeepRangeItems<T>();
	constexpr int kVec = (int)(16 / sizeof(T));
	constexpr int kVecs = kItems / kVec;
	constexpr int kWaves = BLOCK / kWave;
	constexpr int kTile = BLOCK * kItems;
	constexpr int kWaveRows = kItems * kWave;
	static_assert(kVecs * kVec == kItems, "whole vectors per thread");
	__shared__ __align__(16) T s_stage[kTile];
	__shared__ uint32_t s_wtot[kWaves];
	__shared__ uint32_t s_tile;
	__shared__ u64 s_prefix;
	const int tid = threadIdx.x;
	const int lane = tid & (kWave - 1);
	const int wave = tid / kWave;
	if (tid == 0) {
		s_tile = FetchAddAgent(ticket, 1u);
	}
	__syncthreads();
	const uint32_t tile = s_tile;
	const uint64_t wbase = (uint64_t)tile * kTile + (uint64_t)wave * kWaveRows;
	T v[kVecs][kVec];
#pragma unroll
	for (int j = 0; j < kVecs; j++) {
		const uint64_t row0 = wbase + ((uint64_t)j * kWave + lane) * kVec;
		if (row0 + kVec <= n) {
			const uint4 w = *reinterpret_cast<const uint4 *>(in + row0);
			__builtin_memcpy(v[j], &w, sizeof(w));
		} else {
#pragma unroll
			for (int c = 0; c < kVec; c++) {
				v[j][c] = row0 + c < n ? in[row0 + c] : (T)0;
			}
		}
	}
	const u64 lt = LaneMaskLt();
	uint32_t lpos[kVecs];
	uint32_t keep[kVecs];
	uint32_t wrun = 0; // this wave's kept rows so far (the same in every lane)
#pragma unroll
	for (int j = 0; j < kVecs; j++) {
		const uint64_t row0 = wbase + ((uint64_t)j * kWave + lane) * kVec;
		uint32_t k = 0, below = 0, tot = 0;
#pragma unroll
		for (int c = 0; c < kVec; c++) {
			const long long x = (long long)v[j][c];
			const bool ok = row0 + c < n && x >= lo && x <= hi;
			const u64 bal = __ballot(ok);
			k |= ok ? (1u << c) : 0u;
			below += (uint32_t)Popc64(bal & lt);
			tot += (uint32_t)Popc64(bal);
		}
		lpos[j] = wrun + below;
		keep[j] = k;
		wrun += tot;
	}
	if (lane == 0) {
		s_wtot[wave] = wrun;
	}
	__syncthreads();
	uint32_t wexcl = 0, ttot = 0;
#pragma unroll
	for (int w = 0; w < kWaves; w++) {
		const uint32_t c = s_wtot[w];
		wexcl += w < wave ? c : 0u;
		ttot += c;
	}
	if (tid == 0) {
		StoreAgent(state + tile, LbPack(tile == 0 ? kLbInclusive : kLbAggregate, ttot));
	}
	if (wave == 0) {
		const u64 prefix = tile == 0 ? 0ull : LookbackExclusiveWave(state, tile);
		if (lane == 0) {
			if (tile > 0) {
				StoreAgent(state + tile, LbPack(kLbInclusive, prefix + ttot));
			}
			if (tile == ntiles - 1) {
				*count_out = prefix + ttot;
			}
			s_prefix = prefix;
		}
	}
	// Kept rows into LDS in tile order (each lane's earlier components first, then the next lanes').
#pragma unroll
	for (int j = 0; j < kVecs; j++) {
		uint32_t lp = lpos[j] + wexcl;
#pragma unroll
		for (int c = 0; c < kVec; c++) {
			if ((keep[j] >> c) & 1u) {
				s_stage[lp++] = v[j][c];
			}
		}
	}
	__syncthreads();
	const u64 gbase = s_prefix;
	for (uint32_t i = tid; i < ttot; i += BLOCK) {
		out[gbase + i] = s_stage[i];
	}
}

//! Launches the range compaction over n rows (count of kept rows to *count, no host sync): 512-thread tiles
//! (16,384 rows of 4 bytes) from 2^22 rows up, where the per-tile ticket (a same-address atomic, serialized
//! across the device) would otherwise bound the kernel; 256-thread tiles below, so small inputs still fill
//! the device. `state` needs KeepRangeTiles(n) + 1 zeroed words (the last one is the tile ticket).
inline bool KeepRangeWide(uint64_t n) {
	return n >= (1ull << 22);
}
template <typename T>
uint64_t KeepRangeTiles(uint64_t n) {
	const uint64_t tile = (uint64_t)(KeepRangeWide(n) ? 512 : kFgBlock) * KeepRangeItems<T>();
	return (n + tile - 1) / tile;
}
template <typename T>
void LaunchKeepRange(const T *in, uint64_t n, long long lo, long long hi, T *out, u64 *state, u64 *count,
                     hipStream_t stream) {
	const uint64_t ntiles = KeepRangeTiles<T>(n);
	uint32_t *ticket = reinterpret_cast<uint32_t *>(state + ntiles);
	const bool vec = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
	if (vec && KeepRangeWide(n)) {
		hipLaunchKernelGGL((KeepRangeVecKernel<T, 512>), dim3((unsigned)ntiles), dim3(512), 0, stream, in, n, lo,
		                   hi, out, state, ticket, count, ntiles);
	} else if (vec) {
		hipLaunchKernelGGL((KeepRangeVecKernel<T, kFgBlock>), dim3((unsigned)ntiles), dim3(kFgBlock), 0, stream, in,
		                   n, lo, hi, out, state, ticket, count, ntiles);
	} else if (KeepRangeWide(n)) {
		hipLaunchKernelGGL((KeepRangeKernel<T, 512>), dim3((unsigned)ntiles), dim3(512), 0, stream, in, n, lo, hi,
		                   out, state, ticket, count, ntiles);
	} else {
		hipLaunchKernelGGL((KeepRangeKernel<T, kFgBlock>), dim3((unsigned)ntiles), dim3(kFgBlock), 0, stream, in, n,
		                   lo, hi, out, state, ticket, count, ntiles);
	}
	VGPU_HIP(hipGetLastError());
}

template <typename T>
uint64_t KeepRangeT(const void *in, uint64_t n, long long lo, long long hi, void *out, Scratch &scratch,
                    hipStream_t stream) {
	const uint64_t ntiles = KeepRangeTiles<T>(n);
	auto *state = static_cast<u64 *>(scratch.Alloc((ntiles + 2) * sizeof(u64)));
	u64 *count = state + ntiles + 1;
	FillU64(state, 0ull, ntiles + 2, stream);
	{
		ProfileScope prof("keep_key_range", stream, (double)n * sizeof(T));
		LaunchKeepRange<T>((const T *)in, n, lo, hi, (T *)out, state, count, stream);
	}
	uint64_t kept = 0;
	VGPU_HIP(hipMemcpyAsync(&kept, count, 8, hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
	return kept;
}
} // namespace

uint64_t KeepKeyRange(int vtype, const void *in, uint64_t n, long long lo, long long hi, void *out, Scratch &scratch,
                      hipStream_t stream) {
	if (n == 0) {
		return 0;
	}
	switch (vtype) {
	case VT_INT16:
		return KeepRangeT<int16_t>(in, n, lo, hi, out, scratch, stream);
	case VT_INT32:
		return KeepRangeT<int32_t>(in, n, lo, hi, out, scratch, stream);
	case VT_INT64:
		return KeepRangeT<int64_t>(in, n, lo, hi, out, scratch, stream);
	default:
		throw HipError("KeepKeyRange: INT16, INT32 or INT64 keys");
	}
}

void FilterGtI32(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_device,
                 Scratch &scratch, hipStream_t stream) {
	if (n == 0) {
		FillU64(count_device, 0ull, 1, stream);
		return;
	}
	// in[i] > threshold is the signed range [threshold + 1, INT32_MAX] (empty when threshold is INT32_MAX):
	// the key-range compaction kernel: wave-striped loads, contiguous per-wave writes, 16,384-row tiles on
	// large inputs.
	const uint64_t ntiles = KeepRangeTiles<int32_t>(n);
	auto *state = static_cast<u64 *>(scratch.Alloc((ntiles + 1) * sizeof(u64)));
	FillU64(state, 0ull, ntiles + 1, stream);
	ProfileScope prof("filter_gt_i32", stream, (double)n * 4.0);
	LaunchKeepRange<int32_t>(in, n, (long long)threshold + 1, (long long)INT32_MAX, out, state, (u64 *)count_device,
	                         stream);
}

// ------------------------------------------------------------------ FilterGreaterThanInt32 without allocation
//
// The range-compaction kernel needs one look-back word per tile, the tile ticket and the kept count. For
// callers that cannot allocate, they live in `scratch`, one array per GPU for the life of the process:
// kNoAllocMaxTiles + 2 words (1 MiB + 16 B), enough for 2^31 rows at 16,384 rows per tile. In host code
// `scratch` names a host-side stand-in, so the device address of each GPU's copy is resolved once per device
// with hipGetSymbolAddress and only that address reaches kernels and fills.

namespace {
constexpr uint64_t kNoAllocMaxTiles = 1ull << 17;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) u64 scratch[kNoAllocMaxTiles + 2];

std::mutex g_noalloc_mutex;
u64 *g_noalloc_scratch[kNoAllocMaxDevices] = {};
} // namespace

void NoAllocScratchInit() {
	int device = 0;
	VGPU_HIP(hipGetDevice(&device));
	if (device < 0 || device >= kNoAllocMaxDevices) {
		throw HipError("no-allocation scratch: device ordinal out of range");
	}
	std::lock_guard<std::mutex> lock(g_noalloc_mutex);
	if (g_noalloc_scratch[device] != nullptr) {
		return;
	}
	void *ptr = nullptr;
	VGPU_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
	if (ptr == nullptr) {
		throw HipError("no-allocation scratch: hipGetSymbolAddress returned no address");
	}
	g_noalloc_scratch[device] = static_cast<u64 *>(ptr);
}

uint64_t FilterGtI32NoAllocMaxRows() {
	return kNoAllocMaxTiles * 512ull * KeepRangeItems<int32_t>();
}

void FilterGtI32NoAlloc(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_host,
                        hipStream_t stream) {
	if (n > FilterGtI32NoAllocMaxRows()) {
		throw HipError("FilterGreaterThanInt32 (no allocation): " + std::to_string(n) +
		               " rows exceed the fixed scratch limit of " + std::to_string(FilterGtI32NoAllocMaxRows()) +
		               " rows");
	}
	if (n == 0) {
		*count_host = 0;
		return;
	}
	NoAllocScratchInit();
	int device = 0;
	VGPU_HIP(hipGetDevice(&device));
	u64 *state = nullptr;
	{
		std::lock_guard<std::mutex> lock(g_noalloc_mutex);
		state = g_noalloc_scratch[device];
	}
	const uint64_t ntiles = KeepRangeTiles<int32_t>(n);
	// look-back words and the ticket start at zero (a fill kernel, not a memset node, so the call can be
	// captured); the count word is written by the last tile
	FillU64(state, 0ull, ntiles + 1, stream);
	LaunchKeepRange<int32_t>(in, n, (long long)threshold + 1, (long long)INT32_MAX, out, state, state + ntiles + 1,
	                         stream);
	VGPU_HIP(hipMemcpyAsync(count_host, state + ntiles + 1, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
	VGPU_HIP(hipStreamSynchronize(stream));
}

} // namespace vgpu
