// common.h -- device-side building blocks shared by every vectis HIP kernel (gfx950, wave64).
//
// Everything here is written for CDNA4: 64-lane wavefronts, 64-bit ballots, agent-scope atomics for the
// cross-XCD decoupled look-back, and LDS staging. Nothing is shaped for 32-lane warps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgpu {

constexpr int kWave = 64;

typedef uint64_t u64;
typedef int64_t i64;

// ---------------------------------------------------------------- lane helpers

__device__ __forceinline__ int LaneId() {
	return __lane_id();
}

//! Bits of lanes strictly below this lane, as a 64-bit mask.
__device__ __forceinline__ u64 LaneMaskLt() {
	const int lane = LaneId();
	return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ int Popc64(u64 v) {
	return __popcll(v);
}

//! Set bits of `mask` in lanes strictly below this lane (popcount of mask & LaneMaskLt(), two mbcnt steps).
__device__ __forceinline__ uint32_t CountBelow(u64 mask) {
	return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

//! One step of DigitPeers for digit bit B: y = bit B of d sign-extended (all ones where set), the wave's ballot
//! of it, and lo/hi &= ~(bal ^ y), the lanes agreeing with this one on bit B (one v_bitop3 per half, truth
//! table 0x90 over mask, bal, y). The field extract is written as the instruction itself: left to the
//! compiler, y != 0 becomes a shift and a sign test and y a second shift, one more operation per bit.
template <int B>
__device__ __forceinline__ void DigitPeerStep(uint32_t d, uint32_t &lo, uint32_t &hi) {
	uint32_t y;
	asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(y) : "v"(d), "n"(B));
	const u64 bal = __ballot(y != 0u);
	lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)bal, y, 0x90);
	hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(bal >> 32), y, 0x90);
}

template <int B, int R>
__device__ __forceinline__ void DigitPeerSteps(uint32_t d, uint32_t &lo, uint32_t &hi) {
	if constexpr (B < R) {
		DigitPeerStep<B>(d, lo, hi);
		DigitPeerSteps<B + 1, R>(d, lo, hi);
	}
}

//! Lanes of `active` whose R-bit digit equals this lane's (wave multisplit: one ballot per digit bit). Digits
//! are masked to their width, so above a narrower digit every lane's bit is clear, the ballot is empty and
//! the step keeps the mask: all R steps run without a per-bit width test (four VALU operations per bit).
template <int R>
__device__ __forceinline__ u64 DigitPeers(uint32_t d, u64 active) {
	uint32_t lo = (uint32_t)active, hi = (uint32_t)(active >> 32);
	DigitPeerSteps<0, R>(d, lo, hi);
	return ((u64)hi << 32) | lo;
}

// ---------------------------------------------------------------- agent-scope atomics (cross-XCD visible)

__device__ __forceinline__ uint32_t LoadAgent(const uint32_t *p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 LoadAgent(const u64 *p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void StoreAgent(uint32_t *p, uint32_t v) {
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void StoreAgent(u64 *p, u64 v) {
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t FetchAddAgent(uint32_t *p, uint32_t v) {
	return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 FetchAddAgent(u64 *p, u64 v) {
	return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void MinAgent(u64 *p, u64 v) {
	__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void MaxAgent(u64 *p, u64 v) {
	__hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
//! Atomic minimum/maximum issued only when `v` can still change the stored word (the stored word only moves
//! toward the answer, so a value that does not beat a read of it can be dropped). Many waves reducing into
//! one word mostly skip the atomic once the word has settled.
__device__ __forceinline__ void MinAgentIfLower(u64 *p, u64 v) {
	if (v < __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
		MinAgent(p, v);
	}
}
__device__ __forceinline__ void MaxAgentIfHigher(u64 *p, u64 v) {
	if (v > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
		MaxAgent(p, v);
	}
}

// ---------------------------------------------------------------- decoupled look-back words
//
// One 64-bit word per (tile, lane-of-state): [flag:2 | value:62]. A single atomic word carries both the
// status and the value, so a reader never sees a flag without its value, across XCDs, with no fences.

constexpr u64 kLbFlagShift = 62;
constexpr u64 kLbValueMask = (1ull << 62) - 1;
constexpr u64 kLbAggregate = 1ull;
constexpr u64 kLbInclusive = 2ull;

__device__ __forceinline__ u64 LbPack(u64 flag, u64 value) {
	return (flag << kLbFlagShift) | (value & kLbValueMask);
}

//! Exclusive prefix of `tile` over the 64-bit look-back array `state` (one word per tile). A whole wave walks
//! 64 predecessors per step (lane i reads tile-1-i), so a long chain of aggregate-only predecessors costs one
//! round trip per 64 tiles instead of one per tile. Tiles are handed out in launch order by an atomic ticket,
//! so every predecessor is already resident and will publish: the spin terminates.
//! Every lane of the wave must call it; all lanes return the same value.
__device__ __forceinline__ u64 LookbackExclusiveWave(const u64 *state, int64_t tile) {
	const int lane = LaneId();
	u64 exclusive = 0;
	int64_t window_end = tile; // tiles [window_end - 64, window_end) are examined this step
	while (window_end > 0) {
		int64_t t = window_end - 1 - lane;
		// lanes past tile 0 act as an inclusive zero
		u64 word = t >= 0 ? LoadAgent(state + t) : ((u64)kLbInclusive << kLbFlagShift);
		u64 flag = word >> kLbFlagShift;
		// Only the predecessors nearer than the nearest inclusive one are summed, so only they are waited
		// for: one slow tile further back does not hold up the walk.
		u64 inclusive_lanes = __ballot(flag == kLbInclusive);
		int stop_lane = inclusive_lanes ? (__ffsll((long long)inclusive_lanes) - 1) : 64;
		for (;;) {
			const u64 before = stop_lane >= 64 ? ~0ull : ((1ull << stop_lane) - 1);
			if ((__ballot(flag == 0) & before) == 0) {
				break;
			}
			__builtin_amdgcn_s_sleep(1);
			if (flag == 0 && lane < stop_lane) {
				word = LoadAgent(state + t);
				flag = word >> kLbFlagShift;
			}
			inclusive_lanes = __ballot(flag == kLbInclusive);
			stop_lane = inclusive_lanes ? (__ffsll((long long)inclusive_lanes) - 1) : 64;
		}
		u64 value = (t >= 0) ? (word & kLbValueMask) : 0ull;
		u64 contrib = (lane <= stop_lane) ? value : 0ull;
		// wave sum
		for (int offset = 32; offset >= 1; offset >>= 1) {
			contrib += __shfl_xor(contrib, offset);
		}
		exclusive += contrib;
		if (inclusive_lanes) {
			break;
		}
		window_end -= 64;
	}
	return exclusive;
}

// ---------------------------------------------------------------- block scan (exclusive, u32/u64)

//! Exclusive block-wide scan of one value per thread. `lds` needs (blockDim/64 + 1) slots. Returns this
//! thread's exclusive prefix; *total receives the block total (same on every thread).
template <typename T, int BLOCK>
__device__ __forceinline__ T BlockExclusiveScan(T value, T *lds, T *total) {
	constexpr int kWaves = BLOCK / kWave;
	const int lane = LaneId();
	const int wave = threadIdx.x / kWave;
	T inclusive = value;
	for (int offset = 1; offset < kWave; offset <<= 1) {
		T up = __shfl_up(inclusive, offset);
		if (lane >= offset) {
			inclusive += up;
		}
	}
	if (lane == kWave - 1) {
		lds[wave] = inclusive;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		T running = 0;
		for (int w = 0; w < kWaves; w++) {
			T v = lds[w];
			lds[w] = running;
			running += v;
		}
		lds[kWaves] = running;
	}
	__syncthreads();
	T result = lds[wave] + inclusive - value;
	*total = lds[kWaves];
	__syncthreads();
	return result;
}

// ---------------------------------------------------------------- key transforms

//! Order-preserving map of a signed/unsigned fixed-width integer (raw bits in `bits`, width `width` bytes)
//! to uint64: sign-extend then flip the sign bit, so unsigned order equals signed order.
__device__ __host__ __forceinline__ u64 OrderedKey(u64 bits, int width, bool is_signed) {
	if (!is_signed) {
		return bits;
	}
	int shift = 64 - 8 * width;
	int64_t sext = (int64_t)(bits << shift) >> shift;
	return (u64)sext ^ (1ull << 63);
}

__device__ __host__ __forceinline__ u64 UnorderedKey(u64 ordered, int width, bool is_signed) {
	if (!is_signed) {
		return ordered;
	}
	u64 sext = ordered ^ (1ull << 63);
	if (width == 8) {
		return sext;
	}
	return sext & ((1ull << (8 * width)) - 1);
}

__device__ __forceinline__ u64 LoadRawBits(const void *base, uint64_t row, int width) {
	switch (width) {
	case 1:
		return ((const uint8_t *)base)[row];
	case 2:
		return ((const uint16_t *)base)[row];
	case 4:
		return ((const uint32_t *)base)[row];
	default:
		return ((const u64 *)base)[row];
	}
}

// ---------------------------------------------------------------- fill

template <typename T>
__global__ void FillKernel(T *dst, T value, uint64_t n) {
	uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		dst[i] = value;
	}
}

} // namespace vgpu
