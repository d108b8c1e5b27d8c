// onesweep_kernel.h -- the onesweep digit pass kernel and its launcher template (included only by the
// onesweep_*.hip instantiation units).
#pragma once

#include "onesweep.h"

namespace vgpu {
namespace sortpass {

//! One digit pass. Template: R = digit table width (8 or 11 bits), KeyT = compact key (u32/u64), FIRST =
//! the pass packs the raw key columns, NP = payload columns prefetched into registers (0..4), BLOCK = threads
//! per tile (256, or 512 for an 8192-row tile at 16 items per thread: the long per-digit runs of 32 items
//! per thread with half the registers per thread, so twice the waves stay resident).
//!
//! Tile = BLOCK threads x ITEMS items, wave-striped (lane-contiguous 64-item rows, so every global load is a
//! coalesced 64 x width-byte line). Payload values are loaded at kernel entry so their latency overlaps the
//! ranking and the look-back. One LDS buffer (TILE x 4 bytes per round) stages first the per-wave digit
//! counters, then the locally sorted keys, then each payload column 32 bits at a time, so the scatter to
//! global memory writes runs of consecutive addresses per digit bucket.
template <int R, typename KeyT, bool FIRST, int NP, int ITEMS, bool PF, int BLOCK>
__global__ void __launch_bounds__(BLOCK) OnesweepKernel(KeyPack pack, const KeyT *keys_in, KeyT *keys_out,
                                                             PayloadArgs pay, uint64_t n, int digit_shift,
                                                             int digit_bits, const uint32_t *ghist,
                                                             uint32_t *tile_state, uint32_t *tile_ticket,
                                                             const uint64_t *n_dev) {
	constexpr int RADIX = 1 << R;
	constexpr int kWaves = BLOCK / kWave;
	constexpr int TILE = BLOCK * ITEMS;
	constexpr int DPT = (RADIX + BLOCK - 1) / BLOCK;
	constexpr int LDS_WORDS_KEYS = TILE * (int)(sizeof(KeyT) / 4);
	constexpr int LDS_WORDS_COUNT = kWaves * RADIX;
	constexpr int LDS_WORDS_PAY = NP > 0 ? 2 * TILE : 0; // 8-byte payload values staged whole
	constexpr int LDS_WORDS_A = LDS_WORDS_KEYS > LDS_WORDS_COUNT ? LDS_WORDS_KEYS : LDS_WORDS_COUNT;
	constexpr int LDS_WORDS = LDS_WORDS_A > LDS_WORDS_PAY ? LDS_WORDS_A : LDS_WORDS_PAY;
	__shared__ __align__(16) uint32_t s_buf[LDS_WORDS];
	__shared__ uint32_t s_digit_base[RADIX];
	__shared__ uint32_t s_scan[kWaves + 1];
	__shared__ uint32_t s_tile;

	if (n_dev != nullptr) {
		// Row count written by an earlier kernel; the grid covers its bound. Blocks past the count leave
		// before taking a ticket, so the tickets handed out are exactly the tiles that exist.
		n = *n_dev;
		if ((uint64_t)blockIdx.x * TILE >= n) {
			return;
		}
	}
	const int tid = threadIdx.x;
	const int lane = tid & (kWave - 1);
	const int wave = tid / kWave;
	if (tid == 0) {
		s_tile = FetchAddAgent(tile_ticket, 1u);
	}
	for (int i = tid; i < kWaves * RADIX; i += BLOCK) {
		s_buf[i] = 0;
	}
	__syncthreads();
	const uint32_t tile = s_tile;
	const uint64_t tile_base = (uint64_t)tile * TILE;
	const uint64_t wave_base = tile_base + (uint64_t)wave * ITEMS * kWave;
	const uint32_t digit_mask = (1u << digit_bits) - 1;

	KeyT key[ITEMS];
	if constexpr (FIRST) {
		if (pack.ncols == 1) {
			// Common case: one key column. Width dispatch once, then 16 straight-line coalesced loads.
			u64 raw[ITEMS];
			LoadWaveItems<ITEMS>(pack.col[0], pack.width[0], wave_base, lane, n, raw);
#pragma unroll
			for (int k = 0; k < ITEMS; k++) {
				u64 ord = OrderedKey(raw[k], pack.width[0], pack.is_signed[0] != 0);
				key[k] = (KeyT)((ord - pack.min[0]) << pack.shift[0]);
			}
		} else {
#pragma unroll
			for (int k = 0; k < ITEMS; k++) {
				uint64_t row = wave_base + (uint64_t)k * kWave + lane;
				key[k] = row < n ? (KeyT)PackRow(pack, row) : (KeyT)0;
			}
		}
	} else {
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			uint64_t row = wave_base + (uint64_t)k * kWave + lane;
			key[k] = row < n ? keys_in[row] : (KeyT)0;
		}
	}
	// Payload prefetch (raw bits), issued before the ranking so its latency is hidden (PF); without PF each
	// column is loaded just before its scatter, which holds fewer registers and raises occupancy.
	u64 pv[(NP > 0 && PF) ? NP : 1][ITEMS];
	auto load_payload = [&](int p, u64(&dst)[ITEMS]) {
		if (FIRST && pay.in[p] == nullptr) {
#pragma unroll
			for (int k = 0; k < ITEMS; k++) {
				dst[k] = wave_base + (uint64_t)k * kWave + lane;
			}
		} else if (FIRST && ((pay.narrow >> p) & 1u) != 0) {
			LoadWaveItemsT<ITEMS>((const u64 *)pay.in[p], wave_base, lane, n, dst);
			bool wide = false;
#pragma unroll
			for (int k = 0; k < ITEMS; k++) {
				wide = wide || (long long)dst[k] != (long long)(int)(uint32_t)dst[k];
			}
			if (wide) {
				atomicOr(pay.overflow, 1u);
			}
		} else {
			LoadWaveItems<ITEMS>(pay.in[p], pay.width[p], wave_base, lane, n, dst);
		}
	};
	if constexpr (PF) {
#pragma unroll
		for (int p = 0; p < NP; p++) {
			if (p < pay.n) {
				load_payload(p, pv[p]);
			}
		}
	}

	// Wave-level multisplit: peers = lanes of this round holding the same digit. Rank = earlier rounds'
	// count of that digit in this wave + earlier peers in this round, which keeps the sort stable.
	uint32_t *my_counts = s_buf + wave * RADIX;
	uint32_t pos[ITEMS];
#pragma unroll
	for (int k = 0; k < ITEMS; k++) {
		uint64_t row = wave_base + (uint64_t)k * kWave + lane;
		bool valid = row < n;
		uint32_t d = (uint32_t)(key[k] >> digit_shift) & digit_mask;
		const u64 peers = DigitPeers<R>(d, __ballot(valid));
		uint32_t r = CountBelow(peers);
		uint32_t base = valid ? my_counts[d] : 0u;
		__builtin_amdgcn_wave_barrier();
		if (valid && r == 0) {
			my_counts[d] = base + (uint32_t)Popc64(peers);
		}
		__builtin_amdgcn_wave_barrier();
		pos[k] = base + r;
	}
	__syncthreads();

	// Per-digit block totals; the counters become each wave's exclusive offset within the digit.
	uint32_t dtotal[DPT];
#pragma unroll
	for (int j = 0; j < DPT; j++) {
		int d = tid * DPT + j;
		dtotal[j] = 0;
		if (d < RADIX) {
			uint32_t run = 0;
#pragma unroll
			for (int w = 0; w < kWaves; w++) {
				uint32_t c = s_buf[w * RADIX + d];
				s_buf[w * RADIX + d] = run;
				run += c;
			}
			dtotal[j] = run;
			StoreAgent(tile_state + (uint64_t)tile * RADIX + d, (tile == 0 ? kStFlagInc : kStFlagAgg) | run);
		}
	}
	uint32_t tsum = 0, gsum = 0;
	uint32_t gh[DPT];
#pragma unroll
	for (int j = 0; j < DPT; j++) {
		int d = tid * DPT + j;
		tsum += dtotal[j];
		gh[j] = d < RADIX ? ghist[d] : 0u;
		gsum += gh[j];
	}
	uint32_t block_total;
	uint32_t texcl = BlockExclusiveScan<uint32_t, BLOCK>(tsum, s_scan, &block_total);
	uint32_t gtotal;
	uint32_t gexcl = BlockExclusiveScan<uint32_t, BLOCK>(gsum, s_scan, &gtotal);

	// Look-back: all of this thread's digits advance together, so their loads are issued in parallel.
	uint32_t excl[DPT];
	{
		uint32_t dig[DPT];
		bool act[DPT];
#pragma unroll
		for (int j = 0; j < DPT; j++) {
			dig[j] = (uint32_t)(tid * DPT + j);
			act[j] = tid * DPT + j < RADIX;
		}
		LookBackDigits<DPT>(tile_state, tile, RADIX, dig, act, excl);
	}
	uint32_t bstart = texcl;
	uint32_t goff = gexcl;
#pragma unroll
	for (int j = 0; j < DPT; j++) {
		int d = tid * DPT + j;
		if (d < RADIX) {
			if (tile > 0) {
				StoreAgent(tile_state + (uint64_t)tile * RADIX + d, kStFlagInc | (excl[j] + dtotal[j]));
			}
			s_digit_base[d] = goff + excl[j] - bstart;
#pragma unroll
			for (int w = 0; w < kWaves; w++) {
				s_buf[w * RADIX + d] += bstart;
			}
		}
		bstart += dtotal[j];
		goff += gh[j];
	}
	__syncthreads();

#pragma unroll
	for (int k = 0; k < ITEMS; k++) {
		uint64_t row = wave_base + (uint64_t)k * kWave + lane;
		if (row < n) {
			uint32_t d = (uint32_t)(key[k] >> digit_shift) & digit_mask;
			pos[k] += my_counts[d];
		}
	}
	__syncthreads();
	KeyT *s_keys = reinterpret_cast<KeyT *>(s_buf);
#pragma unroll
	for (int k = 0; k < ITEMS; k++) {
		uint64_t row = wave_base + (uint64_t)k * kWave + lane;
		if (row < n) {
			s_keys[pos[k]] = key[k];
		}
	}
	__syncthreads();

	const uint32_t tile_count = (uint32_t)((n - tile_base) < (uint64_t)TILE ? (n - tile_base) : (uint64_t)TILE);
	uint32_t dst[ITEMS];
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		uint32_t i = tid + j * BLOCK;
		dst[j] = 0;
		if (i < tile_count) {
			KeyT k2 = s_keys[i];
			uint32_t d = (uint32_t)(k2 >> digit_shift) & digit_mask;
			dst[j] = s_digit_base[d] + i;
			keys_out[dst[j]] = k2;
		}
	}

	// Payload: each column staged through LDS in tile order (8-byte values whole), then scattered.
#pragma unroll
	for (int p = 0; p < NP; p++) {
		if (p < pay.n) {
			const int width = pay.width[p];
			u64 pl[PF ? 1 : ITEMS];
			if constexpr (!PF) {
				load_payload(p, pl);
			}
			auto value = [&](int k) -> u64 {
				if constexpr (PF) {
					return pv[p][k];
				} else {
					return pl[k];
				}
			};
			__syncthreads();
			if (width == 8) {
				u64 *s_pay = reinterpret_cast<u64 *>(s_buf);
#pragma unroll
				for (int k = 0; k < ITEMS; k++) {
					uint64_t row = wave_base + (uint64_t)k * kWave + lane;
					if (row < n) {
						s_pay[pos[k]] = value(k);
					}
				}
				__syncthreads();
				u64 *out = reinterpret_cast<u64 *>(pay.out[p]);
#pragma unroll
				for (int j = 0; j < ITEMS; j++) {
					uint32_t i = tid + j * BLOCK;
					if (i < tile_count) {
						out[dst[j]] = s_pay[i];
					}
				}
			} else {
#pragma unroll
				for (int k = 0; k < ITEMS; k++) {
					uint64_t row = wave_base + (uint64_t)k * kWave + lane;
					if (row < n) {
						s_buf[pos[k]] = (uint32_t)value(k);
					}
				}
				__syncthreads();
				void *out = pay.out[p];
#pragma unroll
				for (int j = 0; j < ITEMS; j++) {
					uint32_t i = tid + j * BLOCK;
					if (i < tile_count) {
						uint32_t v = s_buf[i];
						switch (width) {
						case 4:
							reinterpret_cast<uint32_t *>(out)[dst[j]] = v;
							break;
						case 2:
							reinterpret_cast<uint16_t *>(out)[dst[j]] = (uint16_t)v;
							break;
						default:
							reinterpret_cast<uint8_t *>(out)[dst[j]] = (uint8_t)v;
							break;
						}
					}
				}
			}
		}
	}
}

template <int R, int ITEMS, bool PF, int BLOCK, typename KeyT, int NP>
void LaunchTyped(bool first, const KeyPack &pack, const void *keys_in, void *keys_out, const PayloadArgs &pay,
                 uint64_t n, int shift, int bits, const uint32_t *gh, uint32_t *state, uint32_t *ticket,
                 hipStream_t stream, unsigned tiles, const uint64_t *n_dev) {
	if (first) {
		hipLaunchKernelGGL((OnesweepKernel<R, KeyT, true, NP, ITEMS, PF, BLOCK>), dim3(tiles), dim3(BLOCK), 0, stream,
		                   pack, (const KeyT *)keys_in, (KeyT *)keys_out, pay, n, shift, bits, gh, state, ticket,
		                   n_dev);
	} else {
		hipLaunchKernelGGL((OnesweepKernel<R, KeyT, false, NP, ITEMS, PF, BLOCK>), dim3(tiles), dim3(BLOCK), 0,
		                   stream, pack, (const KeyT *)keys_in, (KeyT *)keys_out, pay, n, shift, bits, gh, state,
		                   ticket, n_dev);
	}
}

template <int R, int ITEMS, bool PF, int BLOCK, typename KeyT>
void LaunchKeyed(bool first, const KeyPack &pack, const void *keys_in, void *keys_out, const PayloadArgs &pay,
                 uint64_t n, int shift, int bits, const uint32_t *gh, uint32_t *state, uint32_t *ticket,
                 hipStream_t stream, unsigned tiles, const uint64_t *n_dev) {
	switch (pay.n) {
	case 0:
		LaunchTyped<R, ITEMS, PF, BLOCK, KeyT, 0>(first, pack, keys_in, keys_out, pay, n, shift, bits, gh, state, ticket,
		                                   stream, tiles, n_dev);
		break;
	case 1:
		LaunchTyped<R, ITEMS, PF, BLOCK, KeyT, 1>(first, pack, keys_in, keys_out, pay, n, shift, bits, gh, state, ticket,
		                                   stream, tiles, n_dev);
		break;
	case 2:
		LaunchTyped<R, ITEMS, PF, BLOCK, KeyT, 2>(first, pack, keys_in, keys_out, pay, n, shift, bits, gh, state, ticket,
		                                   stream, tiles, n_dev);
		break;
	case 3:
	case 4:
		LaunchTyped<R, ITEMS, PF, BLOCK, KeyT, 4>(first, pack, keys_in, keys_out, pay, n, shift, bits, gh, state, ticket,
		                                   stream, tiles, n_dev);
		break;
	default:
		throw HipError("radix sort: at most 4 payload columns per pass");
	}
}

template <int R, int ITEMS, bool PF, int BLOCK>
void LaunchVariant(bool first, bool keys64, const KeyPack &pack, const void *keys_in, void *keys_out,
                   const PayloadArgs &pay, uint64_t n, int shift, int bits, const uint32_t *gh, uint32_t *state,
                   uint32_t *ticket, hipStream_t stream, unsigned tiles, const uint64_t *n_dev) {
	if (keys64) {
		LaunchKeyed<R, ITEMS, PF, BLOCK, u64>(first, pack, keys_in, keys_out, pay, n, shift, bits, gh, state, ticket, stream,
		                               tiles, n_dev);
	} else {
		LaunchKeyed<R, ITEMS, PF, BLOCK, uint32_t>(first, pack, keys_in, keys_out, pay, n, shift, bits, gh, state, ticket,
		                                   stream, tiles, n_dev);
	}
	VGPU_HIP(hipGetLastError());
}

} // namespace sortpass
} // namespace vgpu
