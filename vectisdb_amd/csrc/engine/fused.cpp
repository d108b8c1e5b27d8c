// fused.cpp -- source generator and launcher for the fused filter/projection kernels (see fused.h).
#include "fused.h"

#include "../../../include/kernel_cache.hpp"
#include "../hip/join_kernels.h"
#include "../hip/kernels.h"

#include <algorithm>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace vgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kVec = 4;
constexpr int kRounds = 2;
constexpr int kTile = kBlock * kVec * kRounds; // smallest tile: 2048 rows (or pairs, or selection entries)

//! Rounds of kVec rows per lane in a tile: narrow kernels (at most 4 input columns and 4 outputs) take 4
//! (4096-row tiles: half the per-tile scan, look-back and staging overhead), cross products of at most 4
//! columns 8 (the pairs come from cache, so per-tile overhead is all there is), wide kernels keep 2 so
//! their register tiles stay small.
int FusedRounds(const FusedSpec &s) {
	const bool narrow = s.cols.size() <= 4 && s.projections.size() <= 4;
	if (s.cross) {
		return narrow ? 8 : 4;
	}
	return narrow ? 4 : kRounds;
}

//! Tile rounds of the variant used for large inputs (0: none). A narrow row kernel over many rows runs
//! 8,192-row tiles (half the tickets and look-backs of 4,096-row tiles: 4.70 -> 5.06 TB/s over 2^28
//! rows); below kLargeRows that would leave CUs idle (128 tiles for 2^20 rows), so the 4,096-row kernel
//! stays the default.
int FusedRoundsLarge(const FusedSpec &s) {
	const bool narrow = s.cols.size() <= 4 && s.projections.size() <= 4;
	return (!s.cross && narrow) ? 8 : 0;
}
constexpr uint64_t kLargeRows = 1ull << 22;

const char *CType(int vtype) {
	switch (vtype) {
	case VT_INT16:
		return "short";
	case VT_INT32:
		return "int";
	case VT_INT64:
		return "long long";
	case VT_FLOAT32:
		return "float";
	case VT_FLOAT64:
		return "double";
	case VT_BOOLEAN:
		return "bool";
	default:
		throw std::runtime_error("fused kernel: unsupported column type (dictionary strings and HUGEINT are not "
		                         "usable in a fused expression)");
	}
}

// Helpers compiled into every generated kernel. DuckDB functions that have no device builtin, the wave
// look-back over 64-bit [flag:2 | value:62] words (agent scope: the tiles of one launch run on all 8
// XCDs), and a 64-lane inclusive scan.
const char *kPreamble = R"VGPU(
typedef unsigned long long vg_u64;
__device__ __forceinline__ double __vgpu_radians(double d) { return d * 0.017453292519943295; }
__device__ __forceinline__ double __vgpu_degrees(double r) { return r * 57.29577951308232; }
struct VArgs {
  const void* cols[16];
  const unsigned char* valid[16];
  void* outs[16];
  unsigned char* out_valid[16];
  vg_u64 n;
  vg_u64 m;
  const unsigned int* sel;
  unsigned int* sel_out;
  vg_u64* state;
  vg_u64* count;
  vg_u64 ntiles;
  vg_u64 cap;
  const vg_u64* seg;
  vg_u64 nseg;
  const vg_u64* seg_row;
  const vg_u64* seg_len;
  vg_u64* range;
  const vg_u64* dims;
};
__device__ __forceinline__ vg_u64 vg_load(const vg_u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void vg_store(vg_u64* p, vg_u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ vg_u64 vg_lookback(const vg_u64* state, long long tile) {
  const int lane = __lane_id();
  vg_u64 excl = 0;
  long long end = tile;
  while (end > 0) {
    long long t = end - 1 - lane;
    vg_u64 word = t >= 0 ? vg_load(state + t) : (2ull << 62), flag = word >> 62;
    // wait only for the predecessors nearer than the nearest inclusive one (the only ones summed)
    vg_u64 inc = __ballot(flag == 2);
    int stop = inc ? (__ffsll((long long)inc) - 1) : 64;
    for (;;) {
      vg_u64 before = stop >= 64 ? ~0ull : ((1ull << stop) - 1);
      if ((__ballot(flag == 0) & before) == 0) break;
      __builtin_amdgcn_s_sleep(1);
      if (flag == 0 && lane < stop) { word = vg_load(state + t); flag = word >> 62; }
      inc = __ballot(flag == 2);
      stop = inc ? (__ffsll((long long)inc) - 1) : 64;
    }
    vg_u64 value = t >= 0 ? (word & ((1ull << 62) - 1)) : 0ull;
    vg_u64 c = lane <= stop ? value : 0ull;
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
    excl += c;
    if (inc) break;
    end -= 64;
  }
  return excl;
}
// In-order rank of a lane's kept rows in one round (4 consecutive rows per lane, bit j of k = row j kept):
// below = kept rows of the lower lanes, total = the wave's kept rows. One ballot per row position and mbcnt
// instead of a 6-step shuffle scan of the per-lane counts.
__device__ __forceinline__ void vg_rank4(unsigned int k, unsigned int& below, unsigned int& total) {
  below = 0; total = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const unsigned long long b = __ballot((k >> j) & 1u);
    below += __builtin_amdgcn_mbcnt_hi((unsigned int)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)b, 0u));
    total += (unsigned int)__popcll(b);
  }
}
)VGPU";

void EmitExprScope(std::ostringstream &o, const FusedSpec &s, const FusedExpr &e, int r, const std::string &out_ptr,
                   const std::string &idx_expr, bool predicate) {
	o << "        {\n";
	size_t k = e.slots.size();
	o << "          const void* inputs[" << (k ? k : 1) << "] = {";
	for (size_t i = 0; i < k; i++) {
		o << (i ? ", " : "") << "(const void*)&__v" << e.slots[i] << "[" << r << "][__j]";
	}
	if (k == 0) {
		o << "0";
	}
	o << "};\n";
	o << "          (void)inputs;\n";
	o << "          const vg_u64 row = 0; (void)row;\n";
	o << "          const vg_u64 idx = " << idx_expr << "; (void)idx;\n";
	o << "          void* out = " << out_ptr << "; (void)out;\n";
	o << "          " << e.body << "\n";
	o << "        }\n";
	(void)s;
	(void)predicate;
}

std::string ValidExpr(const FusedSpec &s, const FusedExpr &e, int r) {
	std::string v;
	for (int slot : e.slots) {
		if (s.cols[slot].has_valid) {
			v += (v.empty() ? "" : " && ") + std::string("__m") + std::to_string(slot) + "[" + std::to_string(r) +
			     "][__j]";
		}
	}
	return v;
}

} // namespace

const char *FusedPreamble() {
	return kPreamble;
}

const char *FusedCType(int vtype) {
	return CType(vtype);
}

std::string FusedKernelName(const FusedSpec &s, int rounds_override) {
	const int rounds = rounds_override > 0 ? rounds_override : FusedRounds(s);
	std::string name = "vectis_";
	if (s.cross) {
		name += s.predicates.empty() ? "cross_expand" : "cross_filter";
	} else if (s.ranges) {
		name += s.persistent ? "batched_filter_persistent" : "batched_filter";
	} else if (s.selection_in) {
		name += s.predicates.empty() ? "gather_project" : "gather_filter";
	} else {
		name += s.predicates.empty() ? "project" : "filter_project";
	}
	return name + "_t" + std::to_string(kBlock * kVec * rounds);
}

std::string GenerateFusedSource(const FusedSpec &s, int rounds_override) {
	if (s.cols.size() > (size_t)kFusedMaxCols || s.projections.size() > (size_t)kFusedMaxOuts) {
		throw std::runtime_error("fused kernel: at most 16 input columns and 16 outputs");
	}
	for (auto *list : {&s.predicates, &s.projections}) {
		for (auto &e : *list) {
			for (int slot : e.slots) {
				if (slot < 0 || slot >= (int)s.cols.size()) {
					throw std::runtime_error("fused kernel: expression input slot out of range");
				}
			}
		}
	}
	const bool filter = !s.predicates.empty();
	const int rounds = rounds_override > 0 ? rounds_override : FusedRounds(s);
	const int tile_rows = kBlock * kVec * rounds;
	if (s.persistent && (!s.ranges || !s.range_slots.empty() || rounds_override > 0)) {
		throw std::runtime_error("fused kernel: the persistent form is for row ranges without range columns");
	}
	std::ostringstream o;
	o << kPreamble;
	o << "extern \"C\" __global__ void __launch_bounds__(" << kBlock << ") " << FusedKernelName(s, rounds)
	  << "(VArgs a) {\n";
	o << "  const int tid = threadIdx.x; const int lane = tid & 63; const int wave = tid >> 6; (void)lane;\n";
	// Launch rows and tiles: kernel arguments, or (persistent form) two device words an earlier kernel wrote.
	if (s.persistent) {
		o << "  const vg_u64 __n = a.dims[0]; const vg_u64 __ntiles = a.dims[1];\n";
	} else {
		o << "  const vg_u64 __n = a.n; const vg_u64 __ntiles = a.ntiles;\n";
	}
	o << "  (void)__n; (void)__ntiles;\n";
	// Output staging buffer: each output (and the selection) leaves the block as one contiguous run.
	// Sized for the widest staged value (outputs, selection ids, validity bytes): 4-byte outputs halve it,
	// which lets more blocks share a CU.
	int stage_width = s.emit_selection ? 4 : 1;
	for (auto &e : s.projections) {
		stage_width = std::max(stage_width, TypeWidth(e.out_type));
	}
	o << "  __shared__ vg_u64 s_stage[" << ((uint64_t)tile_rows * stage_width + 7) / 8 << "];\n";
	if (filter && s.persistent) {
		// A fixed grid of resident blocks. Each block takes one rank from the ticket (ranks follow the order in
		// which blocks start), then runs tiles rank, rank + grid, rank + 2 grid, ...: a tile's look-back only
		// waits on lower tiles, whose blocks have all started, so the sweep cannot deadlock even when fewer
		// blocks are resident than launched, and the ticket is taken once per block instead of once per tile.
		o << "  __shared__ unsigned int s_tile; __shared__ vg_u64 s_wtot[4]; __shared__ vg_u64 s_prefix;\n";
		o << "  if (tid == 0) s_tile = __hip_atomic_fetch_add((unsigned int*)(a.state + __ntiles), 1u, "
		     "__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n";
		o << "  __syncthreads();\n";
		o << "  const vg_u64 __rank = s_tile;\n";
		o << "  for (vg_u64 tile = __rank; tile < __ntiles; tile += gridDim.x) {\n";
	} else if (filter) {
		o << "  __shared__ unsigned int s_tile; __shared__ vg_u64 s_wtot[4]; __shared__ vg_u64 s_prefix;\n";
		o << "  if (tid == 0) s_tile = __hip_atomic_fetch_add((unsigned int*)(a.state + __ntiles), 1u, "
		     "__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n";
		o << "  __syncthreads();\n";
		o << "  const vg_u64 tile = s_tile;\n";
	} else {
		o << "  const vg_u64 tile = blockIdx.x;\n";
	}
	o << "  const vg_u64 tile_base = tile * " << tile_rows << "ull;\n";
	bool has_segment = false;
	for (auto &c : s.cols) {
		has_segment = has_segment || c.segment_id;
	}
	if ((has_segment || s.ranges) && (s.cross || s.selection_in)) {
		throw std::runtime_error("fused kernel: a segment-id column or row ranges need plain row order");
	}
	if (s.ranges && !filter) {
		throw std::runtime_error("fused kernel: row ranges need a predicate (the outputs are compacted)");
	}
	// A row of the launch is live when it is inside the input (and, with row ranges, not padding).
	const std::string live = s.ranges ? "__in[__j]" : "(i0 + __j < a.n)";
	for (size_t c = 0; c < s.cols.size(); c++) {
		const char *t = CType(s.cols[c].vtype);
		if (s.cols[c].segment_id && (s.cols[c].vtype != VT_INT32 || s.cols[c].has_valid)) {
			throw std::runtime_error("fused kernel: a segment-id column is a non-null INT32");
		}
		o << "  const " << t << "* __c" << c << " = (const " << t << "*)a.cols[" << c << "]; (void)__c" << c << ";\n";
		o << "  " << t << " __v" << c << "[" << rounds << "][" << kVec << "];\n";
		if (s.cols[c].has_valid) {
			o << "  const unsigned char* __vc" << c << " = a.valid[" << c << "];\n";
			o << "  bool __m" << c << "[" << rounds << "][" << kVec << "];\n";
		}
	}
	const int nrange = (int)s.range_slots.size();
	if (nrange > 4) {
		throw std::runtime_error("fused kernel: at most 4 range columns");
	}
	for (int slot : s.range_slots) {
		if (slot < 0 || slot >= (int)s.cols.size() || !TypeIsInteger(s.cols[slot].vtype) || s.cols[slot].has_valid) {
			throw std::runtime_error("fused kernel: a range column is a non-null integer column slot");
		}
	}
	if (nrange > 0) {
		o << "  vg_u64 __rlo[" << nrange << "], __rnhi[" << nrange << "];\n";
		o << "#pragma unroll\n  for (int __q = 0; __q < " << nrange << "; __q++) { __rlo[__q] = ~0ull; __rnhi[__q] = ~0ull; }\n";
	}
	o << "  vg_u64 __i0[" << rounds << "];\n";
	o << "  vg_u64 __rowid[" << rounds << "][" << kVec << "];\n";
	o << "  unsigned int __keep[" << rounds << "];\n";
	if (s.cross) {
		// Fast path: the right side is a multiple of the tile, so a tile is one left row x contiguous right
		// rows -- one division per block, vector loads on the right, a broadcast on the left.
		o << "  const bool __aligned_cross = (a.m % " << tile_rows << "ull) == 0;\n";
		o << "  const vg_u64 __tli = __aligned_cross ? tile_base / a.m : 0;\n";
		o << "  const vg_u64 __trj = __aligned_cross ? tile_base - __tli * a.m : 0;\n";
	}
	o << "#pragma unroll\n  for (int __r = 0; __r < " << rounds << "; __r++) {\n";
	o << "    const vg_u64 i0 = tile_base + (vg_u64)wave * " << (kVec * rounds * 64) << "ull + (vg_u64)__r * "
	  << (kVec * 64) << "ull + (vg_u64)lane * " << kVec << "ull;\n";
	o << "    __i0[__r] = i0;\n";
	if (has_segment || s.ranges) {
		// Segment of the lane's first row by binary search over the offsets (segment g covers rows
		// [seg[g], seg[g+1]); empty segments are skipped), then forward steps for the next rows.
		o << "    unsigned int __sg[" << kVec << "];\n";
		o << "    {\n";
		o << "      vg_u64 __lo = 0, __hi = a.nseg;\n";
		o << "      while (__lo < __hi) { vg_u64 __mid = (__lo + __hi) >> 1; if (a.seg[__mid + 1] <= i0) __lo = __mid + 1; else __hi = __mid; }\n";
		o << "      vg_u64 __g = __lo;\n";
		o << "#pragma unroll\n      for (int __j = 0; __j < " << kVec << "; __j++) {\n";
		o << "        while (__g + 1 < a.nseg && a.seg[__g + 1] <= i0 + __j) __g++;\n";
		o << "        __sg[__j] = (unsigned int)__g;\n";
		o << "      }\n";
		o << "    }\n";
	}
	if (s.ranges) {
		o << "    bool __in[" << kVec << "];\n";
		o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) {\n";
		o << "      const vg_u64 __g = __sg[__j];\n";
		o << "      const vg_u64 __off = i0 + __j - a.seg[__g];\n";
		o << "      __in[__j] = i0 + __j < __n && __off < a.seg_len[__g];\n";
		o << "      __rowid[__r][__j] = __in[__j] ? a.seg_row[__g] + __off : 0ull;\n";
		o << "    }\n";
		// one segment, first and last row live: the lane's rows are contiguous table rows
		o << "    const bool __full = __in[0] && __in[" << (kVec - 1) << "] && __sg[0] == __sg[" << (kVec - 1) << "];\n";
	} else {
		o << "    const bool __full = i0 + " << (kVec - 1) << " < a.n;\n";
	}
	if (s.cross) {
		o << "    vg_u64 __lrow[" << kVec << "], __rrow[" << kVec << "];\n";
		o << "    if (__aligned_cross) {\n";
		o << "#pragma unroll\n      for (int __j = 0; __j < " << kVec << "; __j++) { __lrow[__j] = __tli; __rrow[__j] = __trj + (i0 - tile_base) + __j; __rowid[__r][__j] = i0 + __j; }\n";
		o << "    } else {\n";
		o << "      vg_u64 __li = i0 / a.m; vg_u64 __rj = i0 - __li * a.m;\n";
		o << "#pragma unroll\n      for (int __j = 0; __j < " << kVec << "; __j++) {\n";
		o << "        __lrow[__j] = __li; __rrow[__j] = __rj; __rowid[__r][__j] = i0 + __j;\n";
		o << "        __rj++; while (__rj >= a.m) { __rj -= a.m; __li++; }\n";
		o << "      }\n";
		o << "    }\n";
	} else if (s.selection_in) {
		o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) {\n";
		o << "      __rowid[__r][__j] = (i0 + __j < a.n) ? (vg_u64)a.sel[i0 + __j] : 0ull;\n";
		o << "    }\n";
	} else if (!s.ranges) {
		o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) __rowid[__r][__j] = i0 + __j;\n";
	}
	// Loads: 128-bit (or 64/32-bit for narrow types) when the lane's 4 rows are contiguous and aligned.
	auto emit_vector_load = [&](const std::string &cn, const std::string &vn, const char *t, const std::string &first,
	                            const std::string &indent, const std::string &guard, const std::string &index) {
		o << indent << "if (__full && ((vg_u64)(" << cn << " + " << first << ") & " << kVec << " * sizeof(" << t
		  << ") - 1) == 0) {\n";
		o << indent << "  const void* __ap = __builtin_assume_aligned((const void*)(" << cn << " + " << first << "), "
		  << kVec << " * sizeof(" << t << "));\n";
		o << indent << "  __builtin_memcpy(&" << vn << "[0], __ap, " << kVec << " * sizeof(" << t << "));\n";
		o << indent << "} else {\n";
		o << indent << "#pragma unroll\n";
		o << indent << "  for (int __j = 0; __j < " << kVec << "; __j++) " << vn << "[__j] = " << guard << " ? " << cn
		  << "[" << index << "] : (" << t << ")0;\n";
		o << indent << "}\n";
	};
	for (size_t c = 0; c < s.cols.size(); c++) {
		const char *t = CType(s.cols[c].vtype);
		std::string cn = "__c" + std::to_string(c);
		std::string vn = "__v" + std::to_string(c) + "[__r]";
		if (s.cols[c].segment_id) {
			o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) " << vn << "[__j] = (int)__sg[__j];\n";
			continue;
		}
		if (s.cross) {
			if (s.cols[c].side == 0) {
				o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) " << vn
				  << "[__j] = (i0 + __j < a.n) ? " << cn << "[__lrow[__j]] : (" << t << ")0;\n";
			} else {
				o << "    if (__aligned_cross) {\n";
				emit_vector_load(cn, vn, t, "__rrow[0]", "      ", "(i0 + __j < a.n)", "__rrow[0] + __j");
				o << "    } else {\n";
				o << "#pragma unroll\n      for (int __j = 0; __j < " << kVec << "; __j++) " << vn
				  << "[__j] = (i0 + __j < a.n) ? " << cn << "[__rrow[__j]] : (" << t << ")0;\n";
				o << "    }\n";
			}
		} else if (s.selection_in) {
			o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) " << vn << "[__j] = (i0 + __j < a.n) ? "
			  << cn << "[__rowid[__r][__j]] : (" << t << ")0;\n";
		} else if (s.ranges) {
			emit_vector_load(cn, vn, t, "__rowid[__r][0]", "    ", live, "__rowid[__r][__j]");
		} else {
			emit_vector_load(cn, vn, t, "i0", "    ", live, "i0 + __j");
		}
		if (s.cols[c].has_valid) {
			std::string mn = "__m" + std::to_string(c) + "[__r]";
			std::string idx = s.cross ? (s.cols[c].side == 0 ? "__lrow[__j]" : "__rrow[__j]") : "__rowid[__r][__j]";
			o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) " << mn << "[__j] = " << live
			  << " ? (__vc" << c << " == 0 || __vc" << c << "[" << idx << "] != 0) : false;\n";
		}
	}
	// Predicates.
	o << "    unsigned int __k = 0;\n";
	o << "#pragma unroll\n    for (int __j = 0; __j < " << kVec << "; __j++) {\n";
	o << "      bool __ok = " << live << ";\n";
	for (size_t p = 0; p < s.predicates.size(); p++) {
		auto &e = s.predicates[p];
		o << "      bool __p" << p << " = false;\n";
		o << "      if (__ok) {\n";
		std::ostringstream body;
		EmitExprScope(body, s, e, -1, "(void*)&__p" + std::to_string(p), "0", true);
		std::string text = body.str();
		size_t pos;
		while ((pos = text.find("[-1]")) != std::string::npos) {
			text.replace(pos, 4, "[__r]");
		}
		o << text;
		std::string v = ValidExpr(s, e, 0);
		while ((pos = v.find("[0]")) != std::string::npos) {
			v.replace(pos, 3, "[__r]");
		}
		o << "      }\n";
		o << "      __ok = __ok && __p" << p << (v.empty() ? "" : " && (" + v + ")") << ";\n";
	}
	o << "      if (__ok) __k |= 1u << __j;\n";
	for (int q = 0; q < nrange; q++) {
		const int slot = s.range_slots[q];
		const std::string v = "__v" + std::to_string(slot) + "[__r][__j]";
		const std::string ord = TypeIsSigned(s.cols[slot].vtype) ? "((vg_u64)(long long)" + v + " ^ (1ull << 63))"
		                                                           : "(vg_u64)" + v;
		o << "      if (__ok) { const vg_u64 __o = " << ord << "; __rlo[" << q << "] = __o < __rlo[" << q
		  << "] ? __o : __rlo[" << q << "]; __rnhi[" << q << "] = ~__o < __rnhi[" << q << "] ? ~__o : __rnhi[" << q
		  << "]; }\n";
	}
	o << "    }\n";
	o << "    __keep[__r] = __k;\n";
	o << "  }\n";

	// Tile-local output slot of every kept entry (__lpos), and the tile's global base (__gbase).
	o << "  unsigned int __lpos0[" << rounds << "];\n";
	if (filter) {
		o << "  unsigned int __rtot[" << rounds << "];\n";
		o << "  unsigned int __wrun = 0;\n";
		o << "#pragma unroll\n  for (int __r = 0; __r < " << rounds << "; __r++) {\n";
		o << "    unsigned int __below, __tot;\n";
		o << "    vg_rank4(__keep[__r], __below, __tot);\n";
		o << "    __lpos0[__r] = __wrun + __below;\n";
		o << "    __rtot[__r] = __tot;\n";
		o << "    __wrun += __tot;\n";
		o << "  }\n";
		o << "  if (lane == 0) s_wtot[wave] = __wrun;\n";
		o << "  __syncthreads();\n";
		o << "  unsigned int __wexcl = 0, __ttot = 0;\n";
		o << "  for (int w = 0; w < 4; w++) { if (w < wave) __wexcl += (unsigned int)s_wtot[w]; __ttot += (unsigned int)s_wtot[w]; }\n";
		o << "#pragma unroll\n  for (int __r = 0; __r < " << rounds << "; __r++) __lpos0[__r] += __wexcl;\n";
		o << "  if (tid == 0) vg_store(a.state + tile, ((tile == 0 ? 2ull : 1ull) << 62) | (vg_u64)__ttot);\n";
		o << "  if (wave == 0) {\n";
		o << "    vg_u64 __pre = tile == 0 ? 0ull : vg_lookback(a.state, (long long)tile);\n";
		o << "    if (lane == 0) {\n";
		o << "      if (tile > 0) vg_store(a.state + tile, (2ull << 62) | (__pre + __ttot));\n";
		o << "      s_prefix = __pre;\n";
		o << "      if (tile == __ntiles - 1) *a.count = __pre + __ttot;\n";
		o << "    }\n";
		o << "  }\n";
		o << "  __syncthreads();\n";
		o << "  const vg_u64 __gbase = s_prefix;\n";
	} else {
		o << "#pragma unroll\n  for (int __r = 0; __r < " << rounds << "; __r++) __lpos0[__r] = (unsigned int)(__i0[__r] - tile_base);\n";
		o << "  const unsigned int __ttot = (unsigned int)((a.n - tile_base) < " << tile_rows << "ull ? (a.n - tile_base) : "
		  << tile_rows << "ull);\n";
		o << "  const vg_u64 __gbase = tile_base;\n";
	}
	// Writes: every output goes through the LDS stage, then out as one coalesced run of __ttot entries.
	auto emit_stage_pass = [&](const std::string &stage_type, const std::string &write_stmt_prefix,
	                           const std::string &out_expr, const std::function<void()> &emit_value) {
		o << "  {\n";
		o << "    " << stage_type << "* __st = (" << stage_type << "*)s_stage;\n";
		o << "#pragma unroll\n    for (int __r = 0; __r < " << rounds << "; __r++) {\n";
		o << "      unsigned int __lp = __lpos0[__r];\n";
		o << "#pragma unroll\n      for (int __j = 0; __j < " << kVec << "; __j++) {\n";
		o << "        if ((__keep[__r] >> __j) & 1u) {\n";
		o << "          const vg_u64 __o = __lp++;\n";
		emit_value();
		o << "        }\n";
		o << "      }\n";
		o << "    }\n";
		o << "    __syncthreads();\n";
		o << "    for (unsigned int __i = tid; __i < __ttot; __i += " << kBlock << ") {\n";
		o << "      if (__gbase + __i < a.cap) " << write_stmt_prefix << out_expr << "[__gbase + __i] = __st[__i];\n";
		o << "    }\n";
		o << "    __syncthreads();\n";
		o << "  }\n";
		(void)write_stmt_prefix;
	};
	if (s.emit_selection) {
		emit_stage_pass("unsigned int", "", "a.sel_out", [&]() {
			o << "          __st[__o] = (unsigned int)__rowid[__r][__j];\n";
		});
	}
	for (size_t e = 0; e < s.projections.size(); e++) {
		auto &ex = s.projections[e];
		const char *ot = CType(ex.out_type);
		std::string out_cast = std::string("((") + ot + "*)a.outs[" + std::to_string(e) + "])";
		emit_stage_pass(ot, "", out_cast, [&]() {
			std::ostringstream body;
			EmitExprScope(body, s, ex, -1, "(void*)__st", "__o", false);
			std::string text = body.str();
			size_t pos;
			while ((pos = text.find("[-1]")) != std::string::npos) {
				text.replace(pos, 4, "[__r]");
			}
			o << text;
		});
		std::string v = ValidExpr(s, ex, 0);
		if (!v.empty()) {
			size_t pos;
			while ((pos = v.find("[0]")) != std::string::npos) {
				v.replace(pos, 3, "[__r]");
			}
			o << "  if (a.out_valid[" << e << "] != 0) {\n";
			emit_stage_pass("unsigned char", "", "a.out_valid[" + std::to_string(e) + "]", [&]() {
				o << "          __st[__o] = (" << v << ") ? 1 : 0;\n";
			});
			o << "  }\n";
		}
	}
	if (filter && s.persistent) {
		o << "  }\n"; // the tile loop (every stage pass ends with a barrier, so LDS is free for the next tile)
	}
	// Range columns: wave minimum of {min, ~max}, then one atomic per wave, skipped when it cannot lower the
	// stored word.
	for (int q = 0; q < nrange; q++) {
		o << "  {\n";
		o << "    vg_u64 __a = __rlo[" << q << "], __b = __rnhi[" << q << "];\n";
		o << "    for (int __s = 32; __s >= 1; __s >>= 1) { vg_u64 __x = __shfl_xor(__a, __s); vg_u64 __y = __shfl_xor(__b, __s); __a = __x < __a ? __x : __a; __b = __y < __b ? __y : __b; }\n";
		o << "    if (lane == 0) {\n";
		o << "      if (__a < vg_load(a.range + " << 2 * q << ")) __hip_atomic_fetch_min(a.range + " << 2 * q
		  << ", __a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n";
		o << "      if (__b < vg_load(a.range + " << 2 * q + 1 << ")) __hip_atomic_fetch_min(a.range + " << 2 * q + 1
		  << ", __b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n";
		o << "    }\n";
		o << "  }\n";
	}
	o << "}\n";
	return o.str();
}

namespace {
//! Pairs per static tile of the aligned cross-product kernels (VGPU_CROSS_TILE: 4096, 8192 or 16384).
int CrossTile() {
	static const int tile = [] {
		const char *v = std::getenv("VGPU_CROSS_TILE");
		const int t = v != nullptr ? std::atoi(v) : 8192;
		return (t == 4096 || t == 16384) ? t : 8192;
	}();
	return tile;
}
//! Left rows per block of the aligned cross-product count and write kernels (VGPU_CROSS_COUNT_ROWS,
//! VGPU_CROSS_WRITE_ROWS: 1 to 64).
int CrossRows(const char *env, int fallback) {
	const char *v = std::getenv(env);
	const int r = v != nullptr ? std::atoi(v) : fallback;
	return r >= 1 && r <= 64 ? r : fallback;
}
//! Bump allocator over the unused tail of a launch's state buffer (for the scan of the tile counts).
struct SpanScratch : Scratch {
	SpanScratch(uint64_t *begin, uint64_t *end) : next(begin), end(end) {
	}
	void *Alloc(size_t bytes) override {
		const size_t words = (bytes + 7) / 8;
		if (next + words > end) {
			throw std::runtime_error("fused kernel: state buffer too small for the tile-count scan");
		}
		void *p = next;
		next += words;
		return p;
	}
	uint64_t *next;
	uint64_t *end;
};
bool CrossTwoPass(const FusedSpec &spec) {
	return spec.cross && !spec.predicates.empty() && spec.range_slots.empty() && !spec.emit_selection &&
	       !spec.selection_in && !spec.ranges;
}
} // namespace

FusedKernel::~FusedKernel() {
	DestroyRtcKernel(rtc_);
	if (rtc_count_ != nullptr) {
		DestroyRtcKernel(rtc_count_);
	}
	if (rtc_write_ != nullptr) {
		DestroyRtcKernel(rtc_write_);
	}
	if (rtc_large_ != nullptr) {
		DestroyRtcKernel(rtc_large_);
	}
}

size_t FusedKernel::StateBytes(uint64_t n) {
	uint64_t tiles = (n + kTile - 1) / kTile;
	return (tiles + 1) * sizeof(uint64_t);
}

namespace {
struct VArgsHost {
	const void *cols[16];
	const unsigned char *valid[16];
	void *outs[16];
	unsigned char *out_valid[16];
	uint64_t n;
	uint64_t m;
	const uint32_t *sel;
	uint32_t *sel_out;
	uint64_t *state;
	uint64_t *count;
	uint64_t ntiles;
	uint64_t cap;
	const uint64_t *seg;
	uint64_t nseg;
	const uint64_t *seg_row;
	const uint64_t *seg_len;
	uint64_t *range;
	const uint64_t *dims;
};
} // namespace

uint64_t FusedKernel::tile_rows() const {
	return (uint64_t)kBlock * kVec * FusedRounds(spec_);
}

void FusedKernel::Launch(const FusedLaunch &args, void *scratch_state, size_t scratch_bytes, hipStream_t stream) const {
	if (!spec_.range_slots.empty()) {
		if (args.range_device == nullptr) {
			throw std::runtime_error("fused kernel: range columns need a device range buffer");
		}
		FillU64(args.range_device, ~0ull, 2 * spec_.range_slots.size(), stream);
	}
	if (args.n == 0) {
		if (!spec_.predicates.empty() && args.count_device != nullptr) {
			FillU64(args.count_device, 0ull, 1, stream);
		}
		return;
	}
	VArgsHost h {};
	for (size_t c = 0; c < spec_.cols.size(); c++) {
		h.cols[c] = args.cols[c];
		h.valid[c] = args.valid != nullptr ? args.valid[c] : nullptr;
	}
	for (size_t e = 0; e < spec_.projections.size(); e++) {
		h.outs[e] = args.outs[e];
		h.out_valid[e] = args.out_valid != nullptr ? args.out_valid[e] : nullptr;
	}
	h.n = args.n;
	h.m = args.right_rows == 0 ? 1 : args.right_rows;
	h.sel = args.sel;
	h.sel_out = args.sel_out;
	int rounds = FusedRounds(spec_);
	const RtcKernel *rtc = rtc_;
	if (spec_.persistent) {
		if (args.dims_device == nullptr || args.count_device == nullptr || scratch_state == nullptr ||
		    scratch_bytes < StateBytes(args.n)) {
			throw std::runtime_error("fused kernel: the persistent form needs device dims, a count slot and state "
			                         "for its row bound");
		}
		if (args.seg_offsets == nullptr || args.nseg == 0 || args.seg_rows == nullptr || args.seg_lens == nullptr) {
			throw std::runtime_error("fused kernel: row ranges need segment offsets, rows and lengths");
		}
		h.ntiles = 0;
		h.cap = args.capacity == 0 ? args.n : args.capacity;
		h.count = args.count_device;
		h.state = static_cast<uint64_t *>(scratch_state);
		h.seg = args.seg_offsets;
		h.nseg = args.nseg;
		h.seg_row = args.seg_rows;
		h.seg_len = args.seg_lens;
		h.dims = args.dims_device;
		void *pargs[] = {&h};
		double bytes = 0;
		for (auto &c : spec_.cols) {
			bytes += (double)args.n * TypeWidth(c.vtype);
		}
		ProfileScope prof("fused_batched_filter", stream, bytes);
		VGPU_HIP(hipModuleLaunchKernel(rtc->function, persistent_grid_, 1, 1, kBlock, 1, 1, 0, stream, pargs, nullptr));
		return;
	}
	if (large_rounds_ > 0 && args.n >= kLargeRows) {
		std::lock_guard<std::mutex> lock(large_mutex_);
		if (rtc_large_ == nullptr) {
			rtc_large_ = CompileRtc(GenerateFusedSource(spec_, large_rounds_), FusedKernelName(spec_, large_rounds_),
			                        spec_.fast_math);
		}
		rtc = rtc_large_;
		rounds = large_rounds_;
	}
	const uint64_t tile_rows = (uint64_t)kBlock * kVec * rounds;
	const uint64_t tiles = (args.n + tile_rows - 1) / tile_rows;
	h.ntiles = tiles;
	h.cap = args.capacity == 0 ? args.n : args.capacity;
	h.count = args.count_device;
	h.seg = args.seg_offsets;
	h.nseg = args.nseg;
	h.seg_row = args.seg_rows;
	h.seg_len = args.seg_lens;
	h.range = args.range_device;
	for (auto &c : spec_.cols) {
		if (c.segment_id && (args.seg_offsets == nullptr || args.nseg == 0)) {
			throw std::runtime_error("fused kernel: a segment-id column needs segment offsets");
		}
	}
	if (spec_.ranges && (args.seg_offsets == nullptr || args.nseg == 0 || args.seg_rows == nullptr ||
	                     args.seg_lens == nullptr)) {
		throw std::runtime_error("fused kernel: row ranges need segment offsets, rows and lengths");
	}
	if (!spec_.predicates.empty()) {
		if (scratch_state == nullptr || scratch_bytes < StateBytes(args.n)) {
			throw std::runtime_error("fused kernel: look-back state buffer too small");
		}
		if (args.count_device == nullptr) {
			throw std::runtime_error("fused kernel: a filtering launch needs a device count slot");
		}
		h.state = static_cast<uint64_t *>(scratch_state);
	}
	const bool aligned_cross = rtc_count_ != nullptr && args.right_rows >= (uint64_t)cross_tile_ &&
	                           args.right_rows % (uint64_t)cross_tile_ == 0;
	if (!spec_.predicates.empty() && !aligned_cross) {
		FillU64(h.state, 0ull, tiles + 1, stream); // look-back words and the tile ticket
	}
	void *kargs[] = {&h};
	double bytes = 0;
	for (auto &c : spec_.cols) {
		bytes += (double)args.n * TypeWidth(c.vtype);
	}
	if (aligned_cross) {
		// count pass (every static tile writes its kept count), scan to offsets in place (+ the total into
		// *count), write pass; profiled as one operator
		const uint64_t ctiles = args.n / (uint64_t)cross_tile_;
		const uint64_t chunks = args.right_rows / (uint64_t)cross_tile_;
		const uint64_t left = args.n / args.right_rows;
		h.ntiles = ctiles;
		const uint64_t count_grid = chunks * ((left + cross_count_rows_ - 1) / cross_count_rows_);
		const uint64_t write_grid = chunks * ((left + cross_write_rows_ - 1) / cross_write_rows_);
		ProfileScope prof("fused_cross_filter", stream, bytes);
		VGPU_HIP(hipModuleLaunchKernel(rtc_count_->function, (unsigned)count_grid, 1, 1, kBlock, 1, 1, 0, stream, kargs,
		                               nullptr));
		SpanScratch span(h.state + ctiles + 1, h.state + scratch_bytes / sizeof(uint64_t));
		ExclusiveScanU64(h.state, h.state, ctiles, args.count_device, span, stream);
		VGPU_HIP(hipModuleLaunchKernel(rtc_write_->function, (unsigned)write_grid, 1, 1, kBlock, 1, 1, 0, stream, kargs,
		                               nullptr));
		return;
	}
	ProfileScope prof(spec_.cross ? "fused_cross_filter" : (spec_.ranges ? "fused_batched_filter" : "fused_filter_project"),
	                  stream, bytes);
	VGPU_HIP(hipModuleLaunchKernel(rtc->function, (unsigned)tiles, 1, 1, kBlock, 1, 1, 0, stream, kargs, nullptr));
}

std::shared_ptr<FusedKernel> GetFusedKernel(const FusedSpec &spec) {
	static std::mutex mutex;
	static std::map<std::string, std::shared_ptr<FusedKernel>> cache;
	std::string source = GenerateFusedSource(spec);
	std::string key = vector_gpu::KernelCache::HashSource(source + (spec.fast_math ? "#fast" : "#strict"));
	std::lock_guard<std::mutex> lock(mutex);
	auto it = cache.find(key);
	if (it != cache.end()) {
		return it->second;
	}
	auto k = std::shared_ptr<FusedKernel>(new FusedKernel());
	k->spec_ = spec;
	k->source = source;
	k->rtc_ = CompileRtc(source, FusedKernelName(spec), spec.fast_math);
	if (CrossTwoPass(spec)) {
		const int tile = CrossTile();
		k->cross_tile_ = tile;
		k->cross_count_rows_ = CrossRows("VGPU_CROSS_COUNT_ROWS", 16);
		k->cross_write_rows_ = CrossRows("VGPU_CROSS_WRITE_ROWS", 2);
		k->rtc_count_ = CompileRtc(GenerateCrossAlignedSource(spec, tile, 1, k->cross_count_rows_),
		                           CrossAlignedKernelName(spec, tile, 1, k->cross_count_rows_), spec.fast_math);
		k->rtc_write_ = CompileRtc(GenerateCrossAlignedSource(spec, tile, 2, k->cross_write_rows_),
		                           CrossAlignedKernelName(spec, tile, 2, k->cross_write_rows_), spec.fast_math);
	}
	k->large_rounds_ = spec.persistent ? 0 : FusedRoundsLarge(spec);
	if (spec.persistent) {
		int per_cu = 0, dev = 0, cus = 0;
		VGPU_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k->rtc_->function, kBlock, 0));
		VGPU_HIP(hipGetDevice(&dev));
		VGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
		k->persistent_grid_ = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, cus);
	}
	cache.emplace(key, k);
	return k;
}

} // namespace vgpu
