// onesweep.h -- shared pieces of the onesweep radix pass (gfx950): constants, payload arguments, key
// packing, wave-striped loads, and the launcher of one kernel variant. The kernel itself lives in
// onesweep_kernel.h and is instantiated per (digit width, items per thread, payload prefetch) variant in its
// own translation unit, so the variants compile in parallel.
#pragma once

#include "common.h"
#include "kernels.h"

namespace vgpu {
namespace sortpass {

constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr uint32_t kStFlagAgg = 1u << 30;
constexpr uint32_t kStFlagInc = 2u << 30;
constexpr uint32_t kStValue = (1u << 30) - 1;

struct PayloadArgs {
	int n = 0;
	const void *in[kMaxPayload] = {};
	void *out[kMaxPayload] = {};
	int width[kMaxPayload] = {};
	//! First pass only: bit p = column p is INT64 read whole and carried as its low 32 bits (width 4); a
	//! value outside the INT32 range sets *overflow, and the caller discards the sort.
	uint32_t narrow = 0;
	uint32_t *overflow = nullptr;
};

struct PassList {
	int npasses = 0;
	int shift[8] = {};
	int bits[8] = {};
};

__device__ __forceinline__ u64 PackRow(const KeyPack &p, uint64_t row) {
	u64 packed = 0;
#pragma unroll
	for (int c = 0; c < kMaxKeyCols; c++) {
		if (c < p.ncols && p.bits[c] > 0) {
			u64 raw = LoadRawBits(p.col[c], row, p.width[c]);
			u64 ord = OrderedKey(raw, p.width[c], p.is_signed[c] != 0);
			packed |= (ord - p.min[c]) << p.shift[c];
		}
	}
	return packed;
}

//! Loads the wave-striped items (row = wave_base + k * 64 + lane) of one column with the width switch
//! hoisted out of the item loop, so all ITEMS loads issue back to back.
template <int ITEMS, typename T>
__device__ __forceinline__ void LoadWaveItemsT(const T *src, uint64_t wave_base, int lane, uint64_t n,
                                               u64 (&out)[ITEMS]) {
#pragma unroll
	for (int k = 0; k < ITEMS; k++) {
		uint64_t row = wave_base + (uint64_t)k * kWave + lane;
		out[k] = row < n ? (u64)src[row] : 0ull;
	}
}

template <int ITEMS>
__device__ __forceinline__ void LoadWaveItems(const void *src, int width, uint64_t wave_base, int lane, uint64_t n,
                                              u64 (&out)[ITEMS]) {
	switch (width) {
	case 8:
		LoadWaveItemsT<ITEMS>((const u64 *)src, wave_base, lane, n, out);
		break;
	case 4:
		LoadWaveItemsT<ITEMS>((const uint32_t *)src, wave_base, lane, n, out);
		break;
	case 2:
		LoadWaveItemsT<ITEMS>((const uint16_t *)src, wave_base, lane, n, out);
		break;
	default:
		LoadWaveItemsT<ITEMS>((const uint8_t *)src, wave_base, lane, n, out);
		break;
	}
}


//! Launches one pass of variant (R, ITEMS, PF, BLOCK) over `tiles` tiles of BLOCK x ITEMS rows; dispatches on
//! the compact key width, FIRST and the payload count (0, 1, 2, or up to 4). With n_dev the kernel reads the
//! row count from device memory (n and `tiles` are then its bound).
template <int R, int ITEMS, bool PF, int BLOCK>
void LaunchVariant(bool first, bool keys64, const KeyPack &pack, const void *keys_in, void *keys_out,
                   const PayloadArgs &pay, uint64_t n, int shift, int bits, const uint32_t *gh, uint32_t *state,
                   uint32_t *ticket, hipStream_t stream, unsigned tiles, const uint64_t *n_dev);

} // namespace sortpass
} // namespace vgpu
