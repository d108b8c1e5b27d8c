// runreduce.hip -- the GROUP BY's last radix pass fused with the reduction (gfx950).
//
// After the LSD passes over every digit but the most significant one, the rows are ordered by the key's low
// bits. The last pass ranks each tile by the top digit in LDS exactly like a sort pass, and since the ranking is
// stable, the tile's rows are then in full-key order: equal keys sit next to each other. Instead of
// scattering 100M rows to be read again by a segmented reduction, each tile reduces its runs of equal keys in
// registers and writes one partial state per run (a "record"). Records land in key order: digit d's records
// fill a region whose size is bounded by the digit's row count and by (distinct low values + tiles), tiles
// find their place inside it by decoupled look-back on per-digit run counts, and within a tile the runs are in
// key order. A run is cut only where a key's rows straddle a tile boundary, so a key with few rows gives one
// record and a heavy key one record per tile it touches. A second, much smaller reduction combines the
// records of each key (the same two-level head scan and tile fix-up as the row reduce) into the groups.
//
// Replaces, for the GROUP BY shapes it supports, the last OnesweepKernel pass plus the head count and the
// segmented reduction over all rows (the reference: cuda_engine/src/operators/gpu_groupby.cu:380-576).
#include "agg_state.h"
#include "common.h"
#include "kernels.h"
#include "onesweep.h"
#include "segreduce_parts.h"

#include <algorithm>
#include <cstdlib>

namespace vgpu {
namespace {

constexpr int kRunRadix = 256;
constexpr int kRunMaxPayload = 4;
//! Records combined per block of the record reduction: 256 threads x 4.
constexpr int kRecBlock = 256;
constexpr int kRecItems = 4;
constexpr uint64_t kRecTile = (uint64_t)kRecBlock * kRecItems;

struct RunPassArgs {
	const void *keys = nullptr; // compact keys ordered by their bits below `shift`
	int npay = 0;
	const void *pay[kRunMaxPayload] = {};
	int width[kRunMaxPayload] = {};
	uint64_t n = 0;
	int shift = 0; // the last digit: bits [shift, shift + bits)
	int bits = 8;
	const uint32_t *ghist = nullptr; // rows per value of the last digit
	uint32_t region_cap = 0;         // records per digit at most: 2^shift + tiles - 1 (saturated)
	uint32_t *tile_state = nullptr;  // tiles x 256 look-back words, zeroed
	uint32_t *ticket = nullptr;      // zeroed
	uint32_t ntiles = 0;
	void *rec_keys = nullptr;        // record keys (KeyT)
	u64 *rec_state = nullptr;        // record states, 2 x components words each
	//! Written by the last tile: [0, 256] dense start of each digit's records (entry 256 = record count),
	//! [257, 513) start of each digit's region in the record arrays.
	uint32_t *layout = nullptr;
	uint64_t *nrec = nullptr;        // the record count
};

__device__ __forceinline__ uint32_t RunSlot(uint32_t p, int items, int stride) {
	return (p % (uint32_t)items) * (uint32_t)stride + p / (uint32_t)items;
}

//! One tile: rank by the last digit (wave multisplit, stable), stage keys in LDS in that order (transposed:
//! sorted row p at slot (p % ITEMS) * STRIDE + p / ITEMS, so thread t reads its ITEMS consecutive sorted rows
//! with conflict-free loads), find the runs of equal keys, stage each payload column the same way, fold each
//! thread's rows serially and stitch runs across threads with a segmented scan, then write each run's key and
//! state at its record position.
template <typename KeyT, int NP, typename SG, int BLOCK, int ITEMS>
__global__ void __launch_bounds__(BLOCK) RunReduceKernel(RunPassArgs a, CompDesc d) {
	static_assert(BLOCK >= kRunRadix, "one thread per digit");
	static_assert(ITEMS <= 32, "head flags are a 32-bit mask");
	constexpr int RADIX = kRunRadix;
	constexpr int kWaves = BLOCK / kWave;
	constexpr int TILE = BLOCK * ITEMS;
	constexpr int STRIDE = BLOCK + 1;
	constexpr int SLOTS = ITEMS * STRIDE;
	constexpr int SC = SG::n;
	constexpr int W_COUNT = kWaves * RADIX;
	constexpr int W_KEYS = SLOTS * (int)(sizeof(KeyT) / 4);
	constexpr int W_PAY = NP > 0 ? 2 * SLOTS : 0;
	constexpr int W_A = W_COUNT > W_KEYS ? W_COUNT : W_KEYS;
	constexpr int W = W_A > W_PAY ? W_A : W_PAY;
	__shared__ __align__(16) uint32_t s_buf[W];
	__shared__ uint32_t s_runs[RADIX];
	__shared__ uint32_t s_rbase[RADIX];
	__shared__ uint32_t s_scan[kWaves + 1];
	__shared__ u64 s_wave_state[kWaves][2 * SC];
	__shared__ int s_wave_flag[kWaves];
	__shared__ uint32_t s_tile;
	__shared__ KeyT s_last_key;

	const int tid = threadIdx.x;
	const int lane = tid & (kWave - 1);
	const int wave = tid / kWave;
	if (tid == 0) {
		s_tile = FetchAddAgent(a.ticket, 1u);
	}
	for (int i = tid; i < W_COUNT; i += BLOCK) {
		s_buf[i] = 0;
	}
	if (tid < RADIX) {
		s_runs[tid] = 0;
	}
	__syncthreads();
	const uint32_t tile = s_tile;
	const uint64_t n = a.n;
	const uint64_t tile_base = (uint64_t)tile * TILE;
	const uint64_t wave_base = tile_base + (uint64_t)wave * ITEMS * kWave;
	const uint32_t dmask = (1u << a.bits) - 1;
	const int dshift = a.shift;
	const KeyT *keys_in = reinterpret_cast<const KeyT *>(a.keys);

	uint32_t pos[ITEMS];
	{
		KeyT key[ITEMS];
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
			key[k] = row < n ? keys_in[row] : (KeyT)0;
		}
		// Wave multisplit, as in the sort pass: rank = this wave's earlier count of the digit + earlier peers.
		uint32_t *my_counts = s_buf + wave * RADIX;
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
			const bool valid = row < n;
			const uint32_t dg = (uint32_t)(key[k] >> dshift) & dmask;
			u64 peers = __ballot(valid);
#pragma unroll
			for (int b = 0; b < 8; b++) {
				if (b < a.bits) {
					const bool bit = (dg >> b) & 1u;
					const u64 bal = __ballot(bit);
					peers &= bit ? bal : ~bal;
				}
			}
			const uint32_t r = (uint32_t)Popc64(peers & LaneMaskLt());
			const uint32_t base = valid ? my_counts[dg] : 0u;
			__builtin_amdgcn_wave_barrier();
			if (valid && r == 0) {
				my_counts[dg] = base + (uint32_t)Popc64(peers);
			}
			__builtin_amdgcn_wave_barrier();
			pos[k] = base + r;
		}
		__syncthreads();
		// Digit totals of the tile; the counters become each wave's start inside the tile's sorted order.
		uint32_t dtotal = 0;
		if (tid < RADIX) {
#pragma unroll
			for (int w = 0; w < kWaves; w++) {
				const uint32_t c = s_buf[w * RADIX + tid];
				s_buf[w * RADIX + tid] = dtotal;
				dtotal += c;
			}
		}
		uint32_t block_total;
		const uint32_t texcl = BlockExclusiveScan<uint32_t, BLOCK>(dtotal, s_scan, &block_total);
		if (tid < RADIX) {
#pragma unroll
			for (int w = 0; w < kWaves; w++) {
				s_buf[w * RADIX + tid] += texcl;
			}
		}
		__syncthreads();
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
			if (row < n) {
				pos[k] += my_counts[(uint32_t)(key[k] >> dshift) & dmask];
			}
		}
		__syncthreads();
		KeyT *s_keys = reinterpret_cast<KeyT *>(s_buf);
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
			if (row < n) {
				s_keys[RunSlot(pos[k], ITEMS, STRIDE)] = key[k];
			}
		}
		__syncthreads();
	}

	// This thread's sorted rows [r0, r0 + ITEMS) and their run heads.
	const uint32_t tile_count = (uint32_t)((n - tile_base) < (uint64_t)TILE ? (n - tile_base) : (uint64_t)TILE);
	const uint32_t r0 = (uint32_t)tid * ITEMS;
	KeyT sk[ITEMS];
	uint32_t heads = 0;
	uint32_t hc = 0;
	KeyT kprev;
	{
		const KeyT *s_keys = reinterpret_cast<const KeyT *>(s_buf);
#pragma unroll
		for (int i = 0; i < ITEMS; i++) {
			sk[i] = r0 + i < tile_count ? s_keys[i * STRIDE + tid] : (KeyT)0;
		}
		kprev = tid > 0 ? s_keys[(ITEMS - 1) * STRIDE + tid - 1] : (KeyT)0;
		if (r0 < tile_count && tile_count <= r0 + ITEMS) {
			s_last_key = sk[tile_count - 1 - r0];
		}
		KeyT before = kprev;
#pragma unroll
		for (int i = 0; i < ITEMS; i++) {
			const bool h = r0 + i < tile_count && (r0 + i == 0 || sk[i] != before);
			before = sk[i];
			if (h) {
				heads |= 1u << i;
				hc++;
				atomicAdd(&s_runs[(uint32_t)(sk[i] >> dshift) & dmask], 1u);
			}
		}
	}
	uint32_t H;
	const uint32_t hb_excl = BlockExclusiveScan<uint32_t, BLOCK>(hc, s_scan, &H);

	// Runs per digit: published at once (aggregate flag) so later tiles can go on; the look-back comes after
	// the payload staging, by when the predecessors have most likely published too.
	const uint32_t rc = tid < RADIX ? s_runs[tid] : 0u;
	if (tid < RADIX) {
		StoreAgent(a.tile_state + (uint64_t)tile * RADIX + tid, (tile == 0 ? sortpass::kStFlagInc : sortpass::kStFlagAgg) | rc);
	}
	uint32_t runs_total;
	const uint32_t rfirst = BlockExclusiveScan<uint32_t, BLOCK>(rc, s_scan, &runs_total);
	const uint32_t rsize = tid < RADIX ? min(a.ghist[tid], a.region_cap) : 0u;
	uint32_t regions_total;
	const uint32_t rstart = BlockExclusiveScan<uint32_t, BLOCK>(rsize, s_scan, &regions_total);

	// Payload columns: staged in sorted order, each thread keeps its ITEMS values (raw bits).
	u64 pv[NP > 0 ? NP : 1][ITEMS];
#pragma unroll
	for (int p = 0; p < NP; p++) {
		if (p < a.npay) {
			u64 v[ITEMS];
			sortpass::LoadWaveItems<ITEMS>(a.pay[p], a.width[p], wave_base, lane, n, v);
			__syncthreads();
			if (a.width[p] == 8) {
				u64 *s_pay = reinterpret_cast<u64 *>(s_buf);
#pragma unroll
				for (int k = 0; k < ITEMS; k++) {
					if (wave_base + (uint64_t)k * kWave + lane < n) {
						s_pay[RunSlot(pos[k], ITEMS, STRIDE)] = v[k];
					}
				}
				__syncthreads();
#pragma unroll
				for (int i = 0; i < ITEMS; i++) {
					pv[p][i] = s_pay[i * STRIDE + tid];
				}
			} else {
#pragma unroll
				for (int k = 0; k < ITEMS; k++) {
					if (wave_base + (uint64_t)k * kWave + lane < n) {
						s_buf[RunSlot(pos[k], ITEMS, STRIDE)] = (uint32_t)v[k];
					}
				}
				__syncthreads();
#pragma unroll
				for (int i = 0; i < ITEMS; i++) {
					pv[p][i] = s_buf[i * STRIDE + tid];
				}
			}
		}
	}

	// Look-back over the predecessors' per-digit run counts (thread d owns digit d).
	if (tid < RADIX) {
		uint32_t excl = 0;
		if (tile > 0) {
			int64_t t = (int64_t)tile - 1;
			while (true) {
				const uint32_t word = LoadAgent(a.tile_state + (uint64_t)t * RADIX + tid);
				const uint32_t flag = word & (3u << 30);
				if (flag == 0) {
					__builtin_amdgcn_s_sleep(1);
					continue;
				}
				excl += word & sortpass::kStValue;
				if (flag == sortpass::kStFlagInc) {
					break;
				}
				t--;
			}
			StoreAgent(a.tile_state + (uint64_t)tile * RADIX + tid, sortpass::kStFlagInc | (excl + rc));
		}
		s_rbase[tid] = rstart + excl - rfirst;
		if (tile == a.ntiles - 1) {
			a.layout[RADIX + 1 + tid] = rstart;
			s_runs[tid] = excl + rc; // this digit's record count (every tile's runs)
		}
	}
	if (tile == a.ntiles - 1) {
		// Dense start of each digit's records, and the record count.
		__syncthreads();
		const uint32_t tot = tid < RADIX ? s_runs[tid] : 0u;
		uint32_t all;
		const uint32_t dstart = BlockExclusiveScan<uint32_t, BLOCK>(tot, s_scan, &all);
		if (tid < RADIX) {
			a.layout[tid] = dstart;
		}
		if (tid == 0) {
			a.layout[RADIX] = all;
			*a.nrec = all;
		}
	}
	__syncthreads();

	KeyT *rec_keys = reinterpret_cast<KeyT *>(a.rec_keys);
	auto write_run = [&](KeyT key, uint32_t run, const CompState &st) {
		const uint64_t at = (uint64_t)s_rbase[(uint32_t)(key >> dshift) & dmask] + run;
		rec_keys[at] = key;
		StateStore<SG>(d, a.rec_state + at * (2 * SC), st);
	};

	// Serial fold of this thread's rows; runs that start and end here are written at once.
	CompState S, lead;
	StateIdentity<SG>(d, S);
	StateIdentity<SG>(d, lead);
	bool seen = false;
	uint32_t cur_run = 0;
	uint32_t nheads = 0;
	KeyT cur_key = (KeyT)0;
	KeyT key_before_first = kprev;
#pragma unroll
	for (int i = 0; i < ITEMS; i++) {
		if (r0 + i < tile_count) {
			if ((heads >> i) & 1u) {
				if (seen) {
					write_run(cur_key, cur_run, S);
				} else {
					lead = S;
					seen = true;
					key_before_first = i > 0 ? sk[i - 1] : kprev;
				}
				cur_run = hb_excl + nheads;
				nheads++;
				cur_key = sk[i];
				StateIdentity<SG>(d, S);
			}
			u64 vals[NP > 0 ? NP : 1];
#pragma unroll
			for (int p = 0; p < (NP > 0 ? NP : 1); p++) {
				vals[p] = (NP > 0 && p < a.npay) ? pv[p][i] : 0ull;
			}
			StateAddRow<(NP > 0 ? NP : 1), SG>(d, S, vals);
		}
	}

	// Segmented inclusive scan of (seen, S) over the block: X covers from the latest head at or before this
	// thread's last row.
	CompState X = S;
	bool F = seen;
	for (int off = 1; off < kWave; off <<= 1) {
		CompState up;
		StateShflUp<SG>(d, up, X, off);
		const bool fup = __shfl_up((int)F, off) != 0;
		if (lane >= off) {
			if (!F) {
				StateCombine<SG>(d, up, X);
				X = up;
			}
			F = F || fup;
		}
	}
	if (lane == kWave - 1) {
		StateStore<SG>(d, s_wave_state[wave], X);
		s_wave_flag[wave] = F ? 1 : 0;
	}
	__syncthreads();
	CompState P;
	StateIdentity<SG>(d, P);
	for (int w = 0; w < wave; w++) {
		CompState t;
		StateLoad<SG>(d, s_wave_state[w], t);
		if (s_wave_flag[w]) {
			P = t;
		} else {
			StateCombine<SG>(d, P, t);
		}
	}
	if (!F) {
		CompState t = P;
		StateCombine<SG>(d, t, X);
		X = t;
	}
	CompState E;
	StateShflUp<SG>(d, E, X, 1);
	if (lane == 0) {
		E = P;
	}
	if (seen && hb_excl > 0) {
		// The run open at this thread's start ends before its first head.
		CompState V = E;
		StateCombine<SG>(d, V, lead);
		write_run(key_before_first, hb_excl - 1, V);
	}
	if (tid == BLOCK - 1) {
		write_run(s_last_key, H - 1, X); // the tile's last run
	}
}

//! Dense record index -> position in the record arrays (digit regions laid end to end, each partly used).
struct RecMap {
	const uint32_t *dstart; // 257 entries in LDS
	const uint32_t *rstart; // 256 entries in LDS
	__device__ __forceinline__ int Digit(uint64_t i) const {
		int lo = 0, hi = kRunRadix; // largest d with dstart[d] <= i
		while (hi - lo > 1) {
			const int mid = (lo + hi) >> 1;
			if ((uint64_t)dstart[mid] <= i) {
				lo = mid;
			} else {
				hi = mid;
			}
		}
		return lo;
	}
};

__device__ __forceinline__ void LoadRecMap(const uint32_t *layout, uint32_t *s_dstart, uint32_t *s_rstart) {
	for (int i = threadIdx.x; i <= kRunRadix; i += blockDim.x) {
		s_dstart[i] = layout[i];
	}
	for (int i = threadIdx.x; i < kRunRadix; i += blockDim.x) {
		s_rstart[i] = layout[kRunRadix + 1 + i];
	}
	__syncthreads();
}

//! Head counts of the record tiles (a record starts a group when its key differs from the previous record's).
template <typename KeyT>
__global__ void __launch_bounds__(kRecBlock) RecHeadsKernel(const KeyT *rec_keys, const uint32_t *layout,
                                                           const uint64_t *nrec, uint32_t *counts) {
	__shared__ uint32_t s_dstart[kRunRadix + 1];
	__shared__ uint32_t s_rstart[kRunRadix];
	__shared__ uint32_t s_scan[kRecBlock / kWave + 1];
	const uint64_t P = *nrec;
	const uint64_t tile_base = (uint64_t)blockIdx.x * kRecTile;
	if (tile_base >= P) {
		return;
	}
	LoadRecMap(layout, s_dstart, s_rstart);
	const RecMap m {s_dstart, s_rstart};
	const uint64_t i0 = tile_base + (uint64_t)threadIdx.x * kRecItems;
	uint32_t hc = 0;
	if (i0 < P) {
		const uint64_t first = i0 > 0 ? i0 - 1 : 0;
		int dg = m.Digit(first);
		KeyT before = (KeyT)0;
		for (uint64_t i = first; i < i0 + kRecItems && i < P; i++) {
			while (i >= s_dstart[dg + 1]) {
				dg++;
			}
			const KeyT k = rec_keys[s_rstart[dg] + (i - s_dstart[dg])];
			if (i >= i0 && (i == 0 || k != before)) {
				hc++;
			}
			before = k;
		}
	}
	uint32_t total;
	(void)BlockExclusiveScan<uint32_t, kRecBlock>(hc, s_scan, &total);
	if (threadIdx.x == 0) {
		counts[blockIdx.x] = total;
	}
}

//! Combines the records of each key into its group: each thread folds 4 consecutive records, runs are stitched
//! across the block with a segmented scan, a group's id is its head's rank (tile prefix of head counts), and a
//! group that continues into the next tile leaves its tail state for the fix-up, as in the row reduce.
template <typename KeyT, typename SG>
__global__ void __launch_bounds__(kRecBlock)
    RecCombineKernel(const KeyT *rec_keys, const u64 *rec_state, const uint32_t *layout, const uint64_t *nrec,
                     CompDesc d, AggOutDesc o, KeyOutDesc ko, const u64 *tile_prefix, const u64 *tile_carry,
                     u64 *lead_state, u64 *tail_state, u64 *tail_gid) {
	constexpr int SC = SG::n;
	constexpr int kWaves = kRecBlock / kWave;
	__shared__ uint32_t s_dstart[kRunRadix + 1];
	__shared__ uint32_t s_rstart[kRunRadix];
	__shared__ u64 s_scan[kWaves + 1];
	__shared__ u64 s_wave_state[kWaves][2 * SC];
	__shared__ int s_wave_flag[kWaves];
	__shared__ KeyT s_tile_last;
	const uint64_t P = *nrec;
	const uint32_t tile = blockIdx.x;
	const uint64_t tile_base = (uint64_t)tile * kRecTile;
	if (tile_base >= P) {
		return;
	}
	LoadRecMap(layout, s_dstart, s_rstart);
	const RecMap m {s_dstart, s_rstart};
	const int tid = threadIdx.x;
	const int lane = tid & (kWave - 1);
	const int wave = tid / kWave;
	const uint64_t i0 = tile_base + (uint64_t)tid * kRecItems;
	const uint64_t tile_end = min(tile_base + kRecTile, P);

	uint64_t at[kRecItems];
	KeyT k[kRecItems];
	KeyT before = (KeyT)0;
	{
		const uint64_t first = i0 > 0 ? i0 - 1 : 0;
		int dg = i0 < P ? m.Digit(first) : 0;
		if (i0 > 0 && i0 < P) {
			before = rec_keys[s_rstart[dg] + (first - s_dstart[dg])];
		}
#pragma unroll
		for (int j = 0; j < kRecItems; j++) {
			const uint64_t i = i0 + j;
			at[j] = 0;
			k[j] = (KeyT)0;
			if (i < P) {
				while (i >= s_dstart[dg + 1]) {
					dg++;
				}
				at[j] = s_rstart[dg] + (i - s_dstart[dg]);
				k[j] = rec_keys[at[j]];
			}
		}
	}
	bool head[kRecItems];
	u64 hc = 0;
	{
		KeyT b = before;
#pragma unroll
		for (int j = 0; j < kRecItems; j++) {
			const uint64_t i = i0 + j;
			head[j] = i < tile_end && (i == 0 || k[j] != b);
			b = k[j];
			hc += head[j] ? 1 : 0;
			if (i + 1 == tile_end) {
				s_tile_last = k[j];
			}
		}
	}
	u64 H;
	const u64 hb_excl = BlockExclusiveScan<u64, kRecBlock>(hc, s_scan, &H);
	const u64 hb_global = tile_prefix[tile] + tile_carry[tile / kScanTiles];

	CompState S, lead;
	StateIdentity<SG>(d, S);
	StateIdentity<SG>(d, lead);
	bool seen = false;
	u64 gid = hb_global + hb_excl;
	u64 cur_gid = 0;
#pragma unroll
	for (int j = 0; j < kRecItems; j++) {
		if (i0 + j < tile_end) {
			if (head[j]) {
				if (seen) {
					WriteGroup<SG>(o, S, cur_gid);
				} else {
					lead = S;
					seen = true;
				}
				cur_gid = gid++;
				WriteKeys(ko, (u64)k[j], cur_gid);
				StateIdentity<SG>(d, S);
			}
			CompState r;
			StateLoad<SG>(d, rec_state + at[j] * (2 * SC), r);
			StateCombine<SG>(d, S, r);
		}
	}
	if (!seen) {
		lead = S;
	}
	CompState X = S;
	bool F = seen;
	for (int off = 1; off < kWave; off <<= 1) {
		CompState up;
		StateShflUp<SG>(d, up, X, off);
		const bool fup = __shfl_up((int)F, off) != 0;
		if (lane >= off) {
			if (!F) {
				StateCombine<SG>(d, up, X);
				X = up;
			}
			F = F || fup;
		}
	}
	if (lane == kWave - 1) {
		StateStore<SG>(d, s_wave_state[wave], X);
		s_wave_flag[wave] = F ? 1 : 0;
	}
	__syncthreads();
	CompState Pst;
	StateIdentity<SG>(d, Pst);
	for (int w = 0; w < wave; w++) {
		CompState t;
		StateLoad<SG>(d, s_wave_state[w], t);
		if (s_wave_flag[w]) {
			Pst = t;
		} else {
			StateCombine<SG>(d, Pst, t);
		}
	}
	if (!F) {
		CompState t = Pst;
		StateCombine<SG>(d, t, X);
		X = t;
	}
	CompState E;
	StateShflUp<SG>(d, E, X, 1);
	if (lane == 0) {
		E = Pst;
	}
	if (seen) {
		CompState V = E;
		StateCombine<SG>(d, V, lead);
		if (hb_excl > 0) {
			WriteGroup<SG>(o, V, hb_global + hb_excl - 1);
		} else {
			StateStore<SG>(d, lead_state + (uint64_t)tile * 2 * SC, V);
		}
	}
	if (tid == kRecBlock - 1) {
		if (H > 0) {
			const u64 gid_last = hb_global + H - 1;
			bool ends = tile_end >= P;
			if (!ends) {
				const int dg = m.Digit(tile_end);
				ends = rec_keys[s_rstart[dg] + (tile_end - s_dstart[dg])] != s_tile_last;
			}
			if (ends) {
				WriteGroup<SG>(o, X, gid_last);
				tail_gid[tile] = ~0ull;
			} else {
				StateStore<SG>(d, tail_state + (uint64_t)tile * 2 * SC, X);
				tail_gid[tile] = gid_last;
			}
		} else {
			StateStore<SG>(d, lead_state + (uint64_t)tile * 2 * SC, X);
			tail_gid[tile] = ~0ull;
		}
	}
}

//! Component signatures the fused pass is compiled for.
using SigCount = StaticSig<SigCode(COMP_COUNT)>;

template <typename KeyT, int NP, typename SG>
void LaunchRunReduce(const RunPassArgs &args, const CompDesc &cd, unsigned tiles, int block, hipStream_t stream) {
	if (block == 1024) {
		hipLaunchKernelGGL((RunReduceKernel<KeyT, NP, SG, 1024, 8>), dim3(tiles), dim3(1024), 0, stream, args, cd);
	} else {
		hipLaunchKernelGGL((RunReduceKernel<KeyT, NP, SG, 512, 16>), dim3(tiles), dim3(512), 0, stream, args, cd);
	}
	VGPU_HIP(hipGetLastError());
}

template <typename KeyT, typename SG>
void LaunchRecordReduce(const void *rec_keys, const u64 *rec_state, const uint32_t *layout, const uint64_t *nrec,
                        uint64_t rec_bound, const CompDesc &cd, const AggOutDesc &od, const KeyOutDesc &ko,
                        u64 *groups_device, Scratch &scratch, hipStream_t stream) {
	const uint64_t ntiles = std::max<uint64_t>((rec_bound + kRecTile - 1) / kRecTile, 1);
	const uint64_t nscan = (ntiles + kScanTiles - 1) / kScanTiles;
	const size_t sw = 2 * (size_t)SG::n;
	auto *tile_heads = static_cast<uint32_t *>(scratch.Alloc(ntiles * sizeof(uint32_t)));
	auto *tile_prefix = static_cast<u64 *>(scratch.Alloc(ntiles * sizeof(u64)));
	auto *tile_carry = static_cast<u64 *>(scratch.Alloc(std::max<uint64_t>(nscan, 1) * sizeof(u64)));
	auto *lead_state = static_cast<u64 *>(scratch.Alloc(ntiles * sw * sizeof(u64)));
	auto *tail_state = static_cast<u64 *>(scratch.Alloc(ntiles * sw * sizeof(u64)));
	auto *tail_gid = static_cast<u64 *>(scratch.Alloc(ntiles * sizeof(u64)));
	auto *run_state = static_cast<u64 *>(scratch.Alloc(ntiles * sw * sizeof(u64)));
	auto *run_open = static_cast<uint8_t *>(scratch.Alloc(ntiles));
	ProfileScope prof("run_records", stream, 0.0);
	hipLaunchKernelGGL((RecHeadsKernel<KeyT>), dim3((unsigned)ntiles), dim3(kRecBlock), 0, stream,
	                   (const KeyT *)rec_keys, layout, nrec, tile_heads);
	VGPU_HIP(hipGetLastError());
	hipLaunchKernelGGL(TilePrefixKernel, dim3((unsigned)nscan), dim3(kScanBlock), 0, stream, tile_heads, ntiles,
	                   tile_prefix, tile_carry, nrec, kRecTile);
	VGPU_HIP(hipGetLastError());
	hipLaunchKernelGGL(TileCarryKernel, dim3(1), dim3(kScanBlock), 0, stream, tile_carry, nscan, groups_device, nrec,
	                   kRecTile);
	VGPU_HIP(hipGetLastError());
	hipLaunchKernelGGL((RecCombineKernel<KeyT, SG>), dim3((unsigned)ntiles), dim3(kRecBlock), 0, stream,
	                   (const KeyT *)rec_keys, rec_state, layout, nrec, cd, od, ko, tile_prefix, tile_carry, lead_state,
	                   tail_state, tail_gid);
	VGPU_HIP(hipGetLastError());
	const uint64_t nchunks = (ntiles + kFixBlock - 1) / kFixBlock;
	hipLaunchKernelGGL(ChunkRunKernel, dim3((unsigned)nchunks), dim3(kFixBlock), 0, stream, cd, tile_heads, lead_state,
	                   run_state, run_open, ntiles, nrec, kRecTile);
	VGPU_HIP(hipGetLastError());
	hipLaunchKernelGGL(TailFinishKernel, dim3((unsigned)nchunks), dim3(kFixBlock), 0, stream, cd, od, run_state,
	                   run_open, tail_state, tail_gid, ntiles, nrec, kRecTile);
	VGPU_HIP(hipGetLastError());
}

//! The signature of a spec the fused pass supports (0 when none): 1 headline mix, 2 its narrowed form, 3 COUNT
//! with an INT32 sum, 4 with an INT64 sum, 5 with a DOUBLE sum, 6 COUNT only.
int RunSig(const CompDesc &cd, int npayload) {
	if (npayload == 2 && SigMatches<SigCountSumMinMaxDd>(cd)) {
		return 1;
	}
	if (npayload == 2 && SigMatches<SigCountSumMinMaxDdN>(cd)) {
		return 2;
	}
	if (npayload == 1 && SigMatches<SigCountSumI32>(cd)) {
		return 3;
	}
	if (npayload == 1 && SigMatches<SigCountSumI64>(cd)) {
		return 4;
	}
	if (npayload == 1 && SigMatches<SigCountSumDd>(cd)) {
		return 5;
	}
	if (npayload == 0 && SigMatches<SigCount>(cd)) {
		return 6;
	}
	return 0;
}

template <typename KeyT, int NP, typename SG>
void RunAll(const RunPassArgs &args, const CompDesc &cd, unsigned tiles, int block, uint64_t rec_bound,
            const AggOutDesc &od, const KeyOutDesc &ko, u64 *groups_device, Scratch &scratch, hipStream_t stream,
            double pass_bytes) {
	{
		ProfileScope prof("radix_pass_reduce", stream, pass_bytes);
		LaunchRunReduce<KeyT, NP, SG>(args, cd, tiles, block, stream);
	}
	LaunchRecordReduce<KeyT, SG>(args.rec_keys, args.rec_state, args.layout, args.nrec, rec_bound, cd, od, ko,
	                             groups_device, scratch, stream);
}

template <typename KeyT>
void RunDispatch(int sig, const RunPassArgs &args, const CompDesc &cd, unsigned tiles, int block,
                 uint64_t rec_bound, const AggOutDesc &od, const KeyOutDesc &ko, u64 *groups_device, Scratch &scratch,
                 hipStream_t stream, double pass_bytes) {
	switch (sig) {
	case 1:
		RunAll<KeyT, 2, SigCountSumMinMaxDd>(args, cd, tiles, block, rec_bound, od, ko, groups_device, scratch, stream,
		                                     pass_bytes);
		break;
	case 2:
		RunAll<KeyT, 2, SigCountSumMinMaxDdN>(args, cd, tiles, block, rec_bound, od, ko, groups_device, scratch,
		                                      stream, pass_bytes);
		break;
	case 3:
		RunAll<KeyT, 1, SigCountSumI32>(args, cd, tiles, block, rec_bound, od, ko, groups_device, scratch, stream,
		                                pass_bytes);
		break;
	case 4:
		RunAll<KeyT, 1, SigCountSumI64>(args, cd, tiles, block, rec_bound, od, ko, groups_device, scratch, stream,
		                                pass_bytes);
		break;
	case 5:
		RunAll<KeyT, 1, SigCountSumDd>(args, cd, tiles, block, rec_bound, od, ko, groups_device, scratch, stream,
		                               pass_bytes);
		break;
	default:
		RunAll<KeyT, 0, SigCount>(args, cd, tiles, block, rec_bound, od, ko, groups_device, scratch, stream,
		                          pass_bytes);
		break;
	}
}

int RunBlock() {
	static const int block = [] {
		const char *e = std::getenv("VGPU_RUN_BLOCK");
		return (e != nullptr && std::atoi(e) == 512) ? 512 : 1024;
	}();
	return block;
}

} // namespace

uint64_t RunReduceRecordBound(uint64_t n, int shift, int bits, uint64_t max_packed) {
	const uint64_t tile = 8192; // both kernel shapes use 8192-row tiles
	const uint64_t tiles = (n + tile - 1) / tile;
	const uint64_t top = shift >= 64 ? 0 : (max_packed >> shift);
	const uint64_t digits = std::min<uint64_t>(top + 1, 1ull << bits);
	const uint64_t cap = shift >= 40 ? n : std::min<uint64_t>(n, (1ull << shift) + tiles - 1);
	const long double bound = (long double)digits * (long double)cap;
	return bound >= (long double)n ? n : (uint64_t)bound;
}

bool RunReduceSupported(const ReduceSpec &spec, int npayload, bool keys64) {
	CompDesc cd;
	AggOutDesc od;
	KeyOutDesc ko;
	BuildReduceDescs(spec, cd, od, ko);
	(void)keys64;
	return RunSig(cd, npayload) != 0;
}

void RunReduceLastPass(const SortedRun &run, uint64_t n, const ReduceSpec &spec, uint64_t *groups_device,
                       Scratch &scratch, hipStream_t stream) {
	if (!run.last_pending || run.last_bits > 8 || run.keys == nullptr) {
		throw HipError("run reduce: the sort did not leave its last 8-bit digit pass");
	}
	if (run.npayload > kRunMaxPayload) {
		throw HipError("run reduce: at most 4 payload columns");
	}
	CompDesc cd;
	AggOutDesc od;
	KeyOutDesc ko;
	BuildReduceDescs(spec, cd, od, ko);
	const int sig = RunSig(cd, run.npayload);
	if (sig == 0) {
		throw HipError("run reduce: aggregate mix not compiled for the fused pass");
	}
	if (n == 0) {
		FillU64(groups_device, 0ull, 1, stream);
		return;
	}
	const int block = RunBlock();
	constexpr uint64_t kTileRows = 8192;
	const uint64_t tiles = (n + kTileRows - 1) / kTileRows;
	const uint64_t rec_bound = RunReduceRecordBound(n, run.last_shift, run.last_bits, run.max_packed);
	const size_t kw = run.keys64 ? 8 : 4;
	const int sc = cd.ncomp;
	RunPassArgs a;
	a.keys = run.keys;
	a.npay = run.npayload;
	for (int p = 0; p < run.npayload; p++) {
		a.pay[p] = run.payload[p];
		a.width[p] = run.payload_width[p];
	}
	a.n = n;
	a.shift = run.last_shift;
	a.bits = run.last_bits;
	a.ghist = run.last_hist;
	a.region_cap = run.last_shift >= 31 ? 0xffffffffu
	                                    : (uint32_t)std::min<uint64_t>((1ull << run.last_shift) + tiles - 1, 0xffffffffull);
	a.ntiles = (uint32_t)tiles;
	auto *state = static_cast<uint32_t *>(scratch.Alloc((tiles * kRunRadix + 1) * sizeof(uint32_t)));
	FillU32(state, 0u, tiles * kRunRadix + 1, stream);
	a.tile_state = state;
	a.ticket = state + tiles * kRunRadix;
	a.rec_keys = scratch.Alloc(std::max<uint64_t>(rec_bound, 1) * kw);
	a.rec_state = static_cast<u64 *>(scratch.Alloc(std::max<uint64_t>(rec_bound, 1) * 2 * sc * sizeof(u64)));
	a.layout = static_cast<uint32_t *>(scratch.Alloc((2 * kRunRadix + 1) * sizeof(uint32_t)));
	a.nrec = static_cast<uint64_t *>(scratch.Alloc(sizeof(uint64_t)));
	double pass_bytes = (double)n * kw;
	for (int p = 0; p < run.npayload; p++) {
		pass_bytes += (double)n * run.payload_width[p];
	}
	if (run.keys64) {
		RunDispatch<u64>(sig, a, cd, (unsigned)tiles, block, rec_bound, od, ko, groups_device, scratch, stream,
		                 pass_bytes);
	} else {
		RunDispatch<uint32_t>(sig, a, cd, (unsigned)tiles, block, rec_bound, od, ko, groups_device, scratch, stream,
		                      pass_bytes);
	}
}

} // namespace vgpu
