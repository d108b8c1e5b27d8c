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

//! Predecessor words read per step of the per-digit look-back (compile time: -DVGPU_LOOKBACK_WINDOW=n).
#ifndef VGPU_LOOKBACK_WINDOW
#define VGPU_LOOKBACK_WINDOW 4
#endif
constexpr int kLookBackWindow = VGPU_LOOKBACK_WINDOW;

//! Decoupled look-back of D digits of one tile over per-(tile, digit) words [flag:2 | value:30] (flag 1 =
//! the tile's own count, 2 = inclusive of every earlier tile, 0 = not yet written): excl[j] = digit[j]'s count
//! in every earlier tile (0 for an inactive digit, or for tile 0). Each step reads the next W predecessor
//! words of every digit together, so a walk past predecessors that have published only their own counts
//! costs one round trip per W tiles instead of one per tile; it stops at the first inclusive word and waits
//! at the first unwritten one.
template <int D, int W = kLookBackWindow>
__device__ __forceinline__ void LookBackDigits(const uint32_t *state, uint64_t tile, uint32_t radix,
                                               const uint32_t (&digit)[D], const bool (&active)[D],
                                               uint32_t (&excl)[D]) {
	int64_t t[D];
	bool done[D];
#pragma unroll
	for (int j = 0; j < D; j++) {
		excl[j] = 0;
		t[j] = (int64_t)tile - 1;
		done[j] = !active[j] || tile == 0;
	}
	while (true) {
		bool all = true;
#pragma unroll
		for (int j = 0; j < D; j++) {
			all = all && done[j];
		}
		if (all) {
			break;
		}
		uint32_t w[D][W];
#pragma unroll
		for (int j = 0; j < D; j++) {
#pragma unroll
			for (int q = 0; q < W; q++) {
				// tile 0 publishes an inclusive word, so no walk reads past it; beyond it stands a zero inclusive
				w[j][q] = (!done[j] && t[j] - q >= 0) ? LoadAgent(state + (uint64_t)(t[j] - q) * radix + digit[j])
				                                      : kStFlagInc;
			}
		}
		bool stall = false;
#pragma unroll
		for (int j = 0; j < D; j++) {
			if (done[j]) {
				continue;
			}
			bool go = true;
			int used = 0;
#pragma unroll
			for (int q = 0; q < W; q++) {
				const uint32_t flag = w[j][q] & (3u << 30);
				if (go && flag != 0) {
					excl[j] += w[j][q] & kStValue;
					used++;
					if (flag == kStFlagInc) {
						done[j] = true;
						go = false;
					}
				} else {
					go = false;
				}
			}
			t[j] -= used;
			stall = stall || (!done[j] && used < W); // t[j] is unwritten: read it again
		}
		if (stall) {
			__builtin_amdgcn_s_sleep(1);
		}
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
