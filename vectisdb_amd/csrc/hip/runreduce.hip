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
#include <string>
#include <type_traits>

namespace vgpu {
namespace {

constexpr int kRunRadix = 256;
constexpr int kRunMaxPayload = 4;
//! Records combined per block of the record reduction: 256 threads x 1 (about a million records: one per
//! thread keeps enough blocks in flight; 4 per thread took 96 us for 1.2M records, latency-bound).
constexpr int kRecBlock = 256;
constexpr int kRecItems = 1;
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

//! Components folded in phase PH: those reading payload column PH, and COUNT in phase 0.
template <typename SG>
constexpr unsigned PhaseMask(int ph) {
	unsigned m = 0;
	for (int c = 0; c < SG::n; c++) {
		const int code = SG::code(c);
		if ((code & 15) == COMP_COUNT) {
			m |= ph == 0 ? (1u << c) : 0u;
		} else if ((code >> 8) - 1 == ph) {
			m |= 1u << c;
		}
	}
	return m;
}

constexpr int VTypeWidth(int vtype) {
	return vtype == VT_INT16 ? 2 : (vtype == VT_INT32 || vtype == VT_FLOAT32) ? 4 : vtype == VT_BOOLEAN ? 1 : 8;
}

//! Bytes per value of payload column p as the signature reads it (the input type of its components).
template <typename SG>
constexpr int PayloadWidth(int p) {
	for (int c = 0; c < SG::n; c++) {
		const int code = SG::code(c);
		if ((code & 15) != COMP_COUNT && (code >> 8) - 1 == p) {
			return VTypeWidth((code >> 4) & 15);
		}
	}
	return 8;
}

template <int W>
using ColT = typename std::conditional<W == 8, u64,
                                       typename std::conditional<W == 4, uint32_t,
                                                                 typename std::conditional<W == 2, uint16_t,
                                                                                           uint8_t>::type>::type>::type;

// ---------------------------------------------------------------- tile-local run states
//
// Inside one tile a run has at most 2^16 rows, so its state fits narrower types than a group's: COUNT in 32
// bits, the sum of a column of at most 32 bits in 64, MIN/MAX of such a column in 32. The fold and the
// cross-thread scans run on these (half the instructions and shuffles of the 128-bit group state); a run's
// record is written in the group state's layout (two 64-bit words per component), so the record combine
// is the ordinary state combine.

template <int CODE, bool ACTIVE>
struct TComp {
	// An inactive component (not folded in this phase): no storage, no work.
	__device__ __forceinline__ void identity() {}
	__device__ __forceinline__ void add(u64) {}
	__device__ __forceinline__ void combine(const TComp &) {}
	__device__ __forceinline__ void shfl_up(const TComp &, int) {}
	__device__ __forceinline__ void store(u64 *) const {}
	__device__ __forceinline__ void load(const u64 *) {}
};

constexpr int CodeKind(int code) {
	return code & 15;
}
constexpr int CodeType(int code) {
	return (code >> 4) & 15;
}
constexpr bool CodeNarrowInt(int code) {
	return CodeType(code) == VT_INT32 || CodeType(code) == VT_INT16 || CodeType(code) == VT_BOOLEAN;
}

template <int CODE>
struct TComp<CODE, true> {
	static constexpr int K = CodeKind(CODE);
	static constexpr int T = CodeType(CODE);
	static constexpr bool kWideSum = K == COMP_SUM_I128 && !CodeNarrowInt(CODE);
	static constexpr bool kNarrowMinMax = (K == COMP_MIN_I64 || K == COMP_MAX_I64) && CodeNarrowInt(CODE);
	// COUNT: u32 | SUM of <= 32-bit ints: i64 | SUM of INT64: (lo, hi) | SUM_DD: (hi, lo) doubles |
	// MIN/MAX of <= 32-bit ints: i32 | MIN/MAX of INT64: i64 | MIN/MAX of doubles: double bits
	using A = typename std::conditional<
	    K == COMP_COUNT || kNarrowMinMax, uint32_t, u64>::type;
	static constexpr bool kHasB = K == COMP_SUM_DD || kWideSum;
	A a;
	u64 b;

	__device__ __forceinline__ void identity() {
		b = 0;
		switch (K) {
		case COMP_MIN_I64:
			a = kNarrowMinMax ? (A)0x7fffffffu : (A)0x7fffffffffffffffull;
			break;
		case COMP_MAX_I64:
			a = kNarrowMinMax ? (A)0x80000000u : (A)0x8000000000000000ull;
			break;
		case COMP_MIN_F64:
			a = (A)AsBits(__builtin_nan(""));
			break;
		case COMP_MAX_F64:
			a = (A)AsBits(-__builtin_inf());
			break;
		default:
			a = 0;
			break;
		}
	}
	__device__ __forceinline__ void add(u64 raw) {
		if constexpr (K == COMP_COUNT) {
			a += 1;
		} else if constexpr (K == COMP_SUM_I128) {
			if constexpr (kWideSum) {
				I128AddI64(a, b, (long long)raw);
			} else {
				a = (A)((long long)a + RawAsI64(raw, T));
			}
		} else if constexpr (K == COMP_SUM_DD) {
			double hi = AsDouble(a), lo = AsDouble(b);
			DdAdd(hi, lo, RawAsF64(raw, T));
			a = AsBits(hi);
			b = AsBits(lo);
		} else if constexpr (K == COMP_MIN_I64 || K == COMP_MAX_I64) {
			if constexpr (kNarrowMinMax) {
				const int v = (int)RawAsI64(raw, T);
				a = (A)(K == COMP_MIN_I64 ? min((int)a, v) : max((int)a, v));
			} else {
				const long long v = (long long)raw;
				a = (A)(K == COMP_MIN_I64 ? (v < (long long)a ? v : (long long)a)
				                          : (v > (long long)a ? v : (long long)a));
			}
		} else if constexpr (K == COMP_MIN_F64) {
			const double v = RawAsF64(raw, T);
			if (DuckLess(v, AsDouble(a))) {
				a = AsBits(v);
			}
		} else if constexpr (K == COMP_MAX_F64) {
			const double v = RawAsF64(raw, T);
			if (DuckLess(AsDouble(a), v)) {
				a = AsBits(v);
			}
		}
	}
	//! this = this (earlier rows) combined with o (later rows)
	__device__ __forceinline__ void combine(const TComp &o) {
		if constexpr (K == COMP_COUNT) {
			a += o.a;
		} else if constexpr (K == COMP_SUM_I128) {
			if constexpr (kWideSum) {
				I128AddI128(a, b, o.a, o.b);
			} else {
				a = (A)((long long)a + (long long)o.a);
			}
		} else if constexpr (K == COMP_SUM_DD) {
			double hi = AsDouble(a), lo = AsDouble(b);
			DdAddDd(hi, lo, AsDouble(o.a), AsDouble(o.b));
			a = AsBits(hi);
			b = AsBits(lo);
		} else if constexpr (K == COMP_MIN_I64 || K == COMP_MAX_I64) {
			if constexpr (kNarrowMinMax) {
				a = (A)(K == COMP_MIN_I64 ? min((int)a, (int)o.a) : max((int)a, (int)o.a));
			} else {
				const long long x = (long long)a, y = (long long)o.a;
				a = (A)(K == COMP_MIN_I64 ? (y < x ? y : x) : (y > x ? y : x));
			}
		} else if constexpr (K == COMP_MIN_F64) {
			if (DuckLess(AsDouble(o.a), AsDouble(a))) {
				a = o.a;
			}
		} else if constexpr (K == COMP_MAX_F64) {
			if (DuckLess(AsDouble(a), AsDouble(o.a))) {
				a = o.a;
			}
		}
	}
	__device__ __forceinline__ void shfl_up(const TComp &src, int delta) {
		a = __shfl_up(src.a, delta);
		if constexpr (kHasB) {
			b = __shfl_up(src.b, delta);
		} else {
			b = 0;
		}
	}
	//! The group state's two words (agg_state.h layout).
	__device__ __forceinline__ void store(u64 *w) const {
		if constexpr (K == COMP_SUM_I128 && !kWideSum) {
			w[0] = (u64)a;
			w[1] = (long long)a < 0 ? ~0ull : 0ull;
		} else if constexpr (kNarrowMinMax) {
			w[0] = (u64)(long long)(int)a;
			w[1] = 0;
		} else {
			w[0] = (u64)a;
			w[1] = b;
		}
	}
	//! Raw words for the cross-wave exchange through LDS (not the group layout).
	__device__ __forceinline__ void put(u64 *w) const {
		w[0] = (u64)a;
		w[1] = b;
	}
	__device__ __forceinline__ void get(const u64 *w) {
		a = (A)w[0];
		b = w[1];
	}
};

//! A phase's run state: components I.. of the signature, those selected by MASK active.
template <unsigned MASK, int I, int... C>
struct TList {
	__device__ __forceinline__ void identity() {}
	template <int NV>
	__device__ __forceinline__ void add(const u64 (&)[NV]) {}
	__device__ __forceinline__ void combine(const TList &) {}
	__device__ __forceinline__ void shfl_up(const TList &, int) {}
	__device__ __forceinline__ void store(u64 *) const {}
	__device__ __forceinline__ void put(u64 *) const {}
	__device__ __forceinline__ void get(const u64 *) {}
};

template <unsigned MASK, int I, int C0, int... Cs>
struct TList<MASK, I, C0, Cs...> {
	static constexpr bool kOn = ((MASK >> I) & 1u) != 0;
	static constexpr int kPayload = (C0 >> 8) - 1;
	TComp<C0, kOn> h;
	TList<MASK, I + 1, Cs...> t;
	__device__ __forceinline__ void identity() {
		h.identity();
		t.identity();
	}
	template <int NV>
	__device__ __forceinline__ void add(const u64 (&v)[NV]) {
		h.add(kPayload >= 0 && kPayload < NV ? v[kPayload >= 0 && kPayload < NV ? kPayload : 0] : 0ull);
		t.add(v);
	}
	__device__ __forceinline__ void combine(const TList &o) {
		h.combine(o.h);
		t.combine(o.t);
	}
	__device__ __forceinline__ void shfl_up(const TList &src, int delta) {
		h.shfl_up(src.h, delta);
		t.shfl_up(src.t, delta);
	}
	__device__ __forceinline__ void store(u64 *rec) const {
		if constexpr (kOn) {
			h.store(rec + 2 * I);
		}
		t.store(rec);
	}
	__device__ __forceinline__ void put(u64 *w) const {
		if constexpr (kOn) {
			h.put(w + 2 * I);
		}
		t.put(w);
	}
	__device__ __forceinline__ void get(const u64 *w) {
		if constexpr (kOn) {
			h.get(w + 2 * I);
		}
		t.get(w);
	}
};

template <typename SG>
struct TileStateOf;
template <int... K>
struct TileStateOf<StaticSig<K...>> {
	template <unsigned MASK>
	using State = TList<MASK, 0, K...>;
};

//! Wave-striped column load (row = wave_base + k * 64 + lane) in the column's own type.
template <int ITEMS, typename T>
__device__ __forceinline__ void LoadWaveCol(const T *src, uint64_t wave_base, int lane, uint64_t n, T (&out)[ITEMS]) {
#pragma unroll
	for (int k = 0; k < ITEMS; k++) {
		const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
		out[k] = row < n ? src[row] : (T)0;
	}
}

//! One tile: rank by the last digit (wave multisplit, stable), stage keys in LDS in that order (transposed:
//! sorted row p at slot (p % ITEMS) * STRIDE + p / ITEMS, so thread t reads its ITEMS consecutive sorted rows
//! with conflict-free loads), find the runs of equal keys, then one phase per payload column: stage the column
//! the same way, fold each thread's rows serially into that column's components, stitch runs across threads
//! with a segmented scan (wave shuffles, then one wave scanning the wave totals), and write those components
//! of each run's record. Column p + 1 is loaded while phase p runs.
template <typename KeyT, int NP, typename SG, int BLOCK, int ITEMS, bool PF1>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BLOCK <= 512 ? 4 : 1)))
RunReduceKernel(RunPassArgs a, CompDesc d) {
	static_assert(BLOCK >= kRunRadix, "one thread per digit");
	static_assert(ITEMS <= 32, "head flags are a 32-bit mask");
	static_assert(NP <= 2, "the fused pass is compiled for at most two payload columns");
	constexpr int RADIX = kRunRadix;
	constexpr int kWaves = BLOCK / kWave;
	constexpr int TILE = BLOCK * ITEMS;
	constexpr int STRIDE = BLOCK + 1;
	constexpr int SLOTS = ITEMS * STRIDE;
	constexpr int SC = SG::n;
	constexpr int W_COUNT = kWaves * RADIX;
	constexpr int W_KEYS = SLOTS * (int)(sizeof(KeyT) / 4);
	constexpr int W0 = NP > 0 ? PayloadWidth<SG>(0) : 4;
	constexpr int W1 = NP > 1 ? PayloadWidth<SG>(1) : 4;
	constexpr int W_PAY = (W0 == 8 || W1 == 8) ? 2 * SLOTS : SLOTS;
	constexpr int W_A = W_COUNT > W_KEYS ? W_COUNT : W_KEYS;
	constexpr int W = W_A > W_PAY ? W_A : W_PAY;
	using T0 = ColT<W0>;
	using T1 = ColT<W1>;
	__shared__ __align__(16) uint32_t s_buf[W];
	__shared__ uint32_t s_runs[RADIX];
	__shared__ uint32_t s_rbase[RADIX];
	__shared__ uint32_t s_scan[kWaves + 1];
	__shared__ u64 s_scan64[kWaves + 1];
	__shared__ u64 s_wave_state[kWaves][2 * SC];
	__shared__ u64 s_wave_pref[kWaves][2 * SC];
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

	// Each row's LDS slot in sorted order, two 16-bit slots per register (a slot is below ITEMS * STRIDE).
	static_assert(ITEMS % 2 == 0 && SLOTS < 65536, "16-bit slots");
	uint32_t slots[ITEMS / 2];
	auto slot_of = [&](int k) -> uint32_t { return (slots[k / 2] >> (16 * (k & 1))) & 0xffffu; };
	KeyT key[ITEMS];
#pragma unroll
	for (int k = 0; k < ITEMS; k++) {
		const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
		key[k] = row < n ? keys_in[row] : (KeyT)0;
	}
	// The first payload column is loaded now: its latency hides behind the ranking.
	T0 col0[ITEMS]; // unused (and dropped) without payload
	if constexpr (NP > 0) {
		LoadWaveCol<ITEMS>(reinterpret_cast<const T0 *>(a.pay[0]), wave_base, lane, n, col0);
	}
	{
		uint32_t pos[ITEMS];
		// Wave multisplit, as in the sort pass: rank = this wave's earlier count of the digit + earlier peers.
		uint32_t *my_counts = s_buf + wave * RADIX;
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
			const bool valid = row < n;
			const uint32_t dg = (uint32_t)(key[k] >> dshift) & dmask;
			const u64 peers = DigitPeers<8>(dg, __ballot(valid));
			const uint32_t r = CountBelow(peers);
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
#pragma unroll
		for (int k = 0; k < ITEMS; k += 2) {
			slots[k / 2] = RunSlot(pos[k], ITEMS, STRIDE) | (RunSlot(pos[k + 1], ITEMS, STRIDE) << 16);
		}
		__syncthreads();
		KeyT *s_keys = reinterpret_cast<KeyT *>(s_buf);
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint64_t row = wave_base + (uint64_t)k * kWave + lane;
			if (row < n) {
				s_keys[slot_of(k)] = key[k];
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
#pragma unroll
		for (int i = 0; i < ITEMS; i++) {
			if (r0 + i + 1 == tile_count) {
				s_last_key = sk[i];
			}
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
	// The key of the run this thread's first head closes (the row just before that head).
	KeyT key_before_first = kprev;
	{
		const int fh = heads != 0 ? __builtin_ctz(heads) : 0;
#pragma unroll
		for (int i = 0; i + 1 < ITEMS; i++) {
			if (i + 1 == fh) {
				key_before_first = sk[i];
			}
		}
	}

	// Runs per digit: published at once (aggregate flag) so later tiles can go on; the look-back comes after
	// the first column is staged, by when the predecessors have most likely published too. One scan gives
	// each digit's first run in the tile (high word) and the start of its record region (low word: regions
	// are min(rows of the digit, region_cap) long, and their total is below 2^32).
	const uint32_t rc = tid < RADIX ? s_runs[tid] : 0u;
	if (tid < RADIX) {
		StoreAgent(a.tile_state + (uint64_t)tile * RADIX + tid,
		           (tile == 0 ? sortpass::kStFlagInc : sortpass::kStFlagAgg) | rc);
	}
	const uint32_t rsize = tid < RADIX ? min(a.ghist[tid], a.region_cap) : 0u;
	u64 both_total;
	const u64 both = BlockExclusiveScan<u64, BLOCK>(((u64)rc << 32) | rsize, s_scan64, &both_total);
	const uint32_t rfirst = (uint32_t)(both >> 32);
	const uint32_t rstart = (uint32_t)both;

	KeyT *rec_keys = reinterpret_cast<KeyT *>(a.rec_keys);
	T1 col1[ITEMS]; // unused (and dropped) below two payload columns

	auto phase = [&](auto ph) {
		constexpr int PH = decltype(ph)::value;
		using TP = typename std::conditional<PH == 0, T0, T1>::type;
		constexpr int WP = PH == 0 ? W0 : W1;
		TP vals[ITEMS];
		if constexpr (NP > 0) {
			const TP *src = nullptr;
			if constexpr (PH == 0) {
				src = col0;
			} else {
				if constexpr (!PF1) {
					LoadWaveCol<ITEMS>(reinterpret_cast<const T1 *>(a.pay[1]), wave_base, lane, n, col1);
				}
				src = col1;
			}
			__syncthreads(); // the LDS buffer is free again (keys or the previous column have been read)
			TP *s_col = reinterpret_cast<TP *>(s_buf);
#pragma unroll
			for (int k = 0; k < ITEMS; k++) {
				if (wave_base + (uint64_t)k * kWave + lane < n) {
					s_col[slot_of(k)] = src[k];
				}
			}
			__syncthreads();
#pragma unroll
			for (int i = 0; i < ITEMS; i++) {
				vals[i] = s_col[i * STRIDE + tid];
			}
			(void)WP;
		}
		if constexpr (PH == 0 && NP > 1 && PF1) {
			// The next column's loads fly while this phase folds.
			LoadWaveCol<ITEMS>(reinterpret_cast<const T1 *>(a.pay[1]), wave_base, lane, n, col1);
		}
		if constexpr (PH == 0) {
			// Look-back over the predecessors' per-digit run counts (thread d owns digit d).
			if (tid < RADIX) {
				uint32_t excl = 0;
				if (tile > 0) {
					const uint32_t dig[1] = {(uint32_t)tid};
					const bool act[1] = {true};
					uint32_t ex[1];
					sortpass::LookBackDigits<1>(a.tile_state, tile, RADIX, dig, act, ex);
					excl = ex[0];
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
		}
		using PS = typename TileStateOf<SG>::template State<PhaseMask<SG>(PH)>;
		auto write_run = [&](KeyT k, uint32_t run, const PS &st) {
			const uint64_t at = (uint64_t)s_rbase[(uint32_t)(k >> dshift) & dmask] + run;
			if constexpr (PH == 0) {
				rec_keys[at] = k;
			}
			st.store(a.rec_state + at * (2 * SC));
		};

		// Serial fold of this thread's rows; runs that start and end here are written at once.
		PS S, lead;
		S.identity();
		lead.identity();
		bool seen = false;
		uint32_t cur_run = 0;
		uint32_t nheads = 0;
		KeyT cur_key = (KeyT)0;
#pragma unroll
		for (int i = 0; i < ITEMS; i++) {
			if (r0 + i < tile_count) {
				if ((heads >> i) & 1u) {
					if (seen) {
						write_run(cur_key, cur_run, S);
					} else {
						lead = S;
						seen = true;
					}
					cur_run = hb_excl + nheads;
					nheads++;
					cur_key = sk[i];
					S.identity();
				}
				u64 v[NP > 0 ? NP : 1];
#pragma unroll
				for (int q = 0; q < (NP > 0 ? NP : 1); q++) {
					v[q] = (NP > 0 && q == PH) ? (u64)vals[i] : 0ull;
				}
				S.add(v);
			}
		}

		// Segmented inclusive scan of (seen, S) over the block: X covers from the latest head at or before this
		// thread's last row.
		PS X = S;
		bool F = seen;
		for (int off = 1; off < kWave; off <<= 1) {
			PS up;
			up.shfl_up(X, off);
			const bool fup = __shfl_up((int)F, off) != 0;
			if (lane >= off) {
				if (!F) {
					up.combine(X);
					X = up;
				}
				F = F || fup;
			}
		}
		if (lane == kWave - 1) {
			X.put(s_wave_state[wave]);
			s_wave_flag[wave] = F ? 1 : 0;
		}
		__syncthreads();
		if (wave == 0) {
			// Exclusive segmented scan of the wave totals, lane w for wave w.
			PS Wst;
			bool WF = false;
			if (lane < kWaves) {
				Wst.get(s_wave_state[lane]);
				WF = s_wave_flag[lane] != 0;
			} else {
				Wst.identity();
			}
			for (int off = 1; off < kWaves; off <<= 1) {
				PS up;
				up.shfl_up(Wst, off);
				const bool fup = __shfl_up((int)WF, off) != 0;
				if (lane >= off) {
					if (!WF) {
						up.combine(Wst);
						Wst = up;
					}
					WF = WF || fup;
				}
			}
			PS Ex;
			Ex.shfl_up(Wst, 1);
			if (lane == 0) {
				Ex.identity();
			}
			if (lane < kWaves) {
				Ex.put(s_wave_pref[lane]);
			}
		}
		__syncthreads();
		PS P;
		P.get(s_wave_pref[wave]);
		if (!F) {
			PS t = P;
			t.combine(X);
			X = t;
		}
		PS E;
		E.shfl_up(X, 1);
		if (lane == 0) {
			E = P;
		}
		if (seen && hb_excl > 0) {
			// The run open at this thread's start ends before its first head.
			PS V = E;
			V.combine(lead);
			write_run(key_before_first, hb_excl - 1, V);
		}
		if (tid == BLOCK - 1) {
			write_run(s_last_key, H - 1, X); // the tile's last run
		}
	};
	phase(std::integral_constant<int, 0> {});
	if constexpr (NP > 1) {
		phase(std::integral_constant<int, 1> {});
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

//! Kernel shapes (threads x rows per thread, and whether the second payload column is loaded while the first
//! is folded): 256 x 16 (the default), 256 x 16 loading the second column at its turn, 512 x 16, 1024 x 8.
//! Measured on the 100M-row, 20-bit-key GROUP BY: 256 x 16 held to 128 VGPRs (4 waves per SIMD, 4 blocks per
//! CU, 28 bytes of spill) 0.76 ms; the same at 151 VGPRs (3 blocks per CU) 0.84 ms; 512 x 16 1.08 ms and
//! 1024 x 8 1.02 ms (one block per CU); the first form, with every column's full group state in registers,
//! 1.23 ms. VGPU_RUN_TILE selects one for sweeps.
struct RunShape {
	int block;
	int items;
	bool prefetch;
};

RunShape RunShapeChoice() {
	static const RunShape shape = [] {
		const char *e = std::getenv("VGPU_RUN_TILE");
		const std::string v = e != nullptr ? e : "";
		if (v == "512x16") {
			return RunShape {512, 16, true};
		}
		if (v == "1024x8") {
			return RunShape {1024, 8, true};
		}
		if (v == "256x16n") {
			return RunShape {256, 16, false};
		}
		return RunShape {256, 16, true};
	}();
	return shape;
}

template <typename KeyT, int NP, typename SG>
void LaunchRunReduce(const RunPassArgs &args, const CompDesc &cd, unsigned tiles, RunShape shape, hipStream_t stream) {
	if (shape.block == 512) {
		hipLaunchKernelGGL((RunReduceKernel<KeyT, NP, SG, 512, 16, true>), dim3(tiles), dim3(512), 0, stream, args, cd);
	} else if (shape.block == 1024) {
		hipLaunchKernelGGL((RunReduceKernel<KeyT, NP, SG, 1024, 8, true>), dim3(tiles), dim3(1024), 0, stream, args, cd);
	} else if (!shape.prefetch) {
		hipLaunchKernelGGL((RunReduceKernel<KeyT, NP, SG, 256, 16, false>), dim3(tiles), dim3(256), 0, stream, args, cd);
	} else {
		hipLaunchKernelGGL((RunReduceKernel<KeyT, NP, SG, 256, 16, true>), dim3(tiles), dim3(256), 0, stream, args, cd);
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
void RunAll(const RunPassArgs &args, const CompDesc &cd, unsigned tiles, RunShape shape, uint64_t rec_bound,
            const AggOutDesc &od, const KeyOutDesc &ko, u64 *groups_device, Scratch &scratch, hipStream_t stream,
            double pass_bytes) {
	for (int p = 0; p < NP; p++) {
		if (args.width[p] != PayloadWidth<SG>(p)) {
			throw HipError("run reduce: payload width does not match the aggregate input type");
		}
	}
	{
		ProfileScope prof("radix_pass_reduce", stream, pass_bytes);
		LaunchRunReduce<KeyT, NP, SG>(args, cd, tiles, shape, stream);
	}
	LaunchRecordReduce<KeyT, SG>(args.rec_keys, args.rec_state, args.layout, args.nrec, rec_bound, cd, od, ko,
	                             groups_device, scratch, stream);
}

template <typename KeyT>
void RunDispatch(int sig, const RunPassArgs &args, const CompDesc &cd, unsigned tiles, RunShape block,
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

} // namespace

uint64_t RunReduceRecordBound(uint64_t n, int shift, int bits, uint64_t max_packed) {
	const RunShape shape = RunShapeChoice();
	const uint64_t tile = (uint64_t)shape.block * shape.items;
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

uint64_t RunReduceStateWords(uint64_t n) {
	const RunShape block = RunShapeChoice();
	const uint64_t tile_rows = (uint64_t)block.block * block.items;
	return ((n + tile_rows - 1) / tile_rows) * kRunRadix + 1;
}

void RunReduceLastPass(const SortedRun &run, uint64_t n, const ReduceSpec &spec, uint64_t *groups_device,
                       Scratch &scratch, hipStream_t stream, uint32_t *state) {
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
	const RunShape block = RunShapeChoice();
	const uint64_t tile_rows = (uint64_t)block.block * block.items;
	const uint64_t tiles = (n + tile_rows - 1) / tile_rows;
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
	// Rank on the digit bits the keys can use: the top digit of the largest packed key bounds them (a 20-bit
	// key range leaves 4 of the last pass's 8 bits, half the ballots of the multisplit).
	const uint64_t top = run.last_shift >= 64 ? 0 : (run.max_packed >> run.last_shift);
	int used = 1;
	while (used < run.last_bits && (top >> used) != 0) {
		used++;
	}
	a.bits = used;
	a.ghist = run.last_hist;
	a.region_cap = run.last_shift >= 31 ? 0xffffffffu
	                                    : (uint32_t)std::min<uint64_t>((1ull << run.last_shift) + tiles - 1, 0xffffffffull);
	a.ntiles = (uint32_t)tiles;
	if (state == nullptr) {
		state = static_cast<uint32_t *>(scratch.Alloc((tiles * kRunRadix + 1) * sizeof(uint32_t)));
		FillU32(state, 0u, tiles * kRunRadix + 1, stream);
	}
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
