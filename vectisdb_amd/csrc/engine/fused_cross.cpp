// fused_cross.cpp -- generated count and write kernels of a filtered cross product whose right side is a whole
// number of tiles (see GenerateCrossAlignedSource in fused.h).
//
// Each block keeps one tile-long chunk of the right side in registers and runs it against `rows_per_block`
// consecutive left rows, so the right side is read from cache once per group of left rows, not once per row.
//
// Pair f of the left-major enumeration is (f / m, f % m) (cuda_engine/src/operators/gpu_cross_join.cu:16-19).
// With m a multiple of the tile, tile t holds left row t * tile / m and the contiguous right rows from
// t * tile % m, so a tile's left columns are one value each. The count kernel writes each tile's kept-pair
// count; a one-block scan turns the counts into output offsets; the write kernel evaluates the predicates
// again (its inputs come from cache) and writes the kept pairs of the tile as one run per output.
#include "fused.h"

#include "../hip/kernels.h"

#include <algorithm>
#include <functional>
#include <sstream>
#include <stdexcept>

namespace vgpu {

namespace {

constexpr int kCrossBlock = 256;
constexpr int kCrossVec = 4;

// 16-byte output store (non-temporal stores measured slower: 0.30 against 0.27 ms for 4,096 x 65,536 pairs)
const char *kStore16 = R"VGPU(
__device__ __forceinline__ void vg_store16(void* p, uint4 w) { *(uint4*)__builtin_assume_aligned(p, 16) = w; }
)VGPU";

// Run writers: 16-byte stores in the aligned middle of [g0, g0 + cnt) (indices at or past `cap` dropped).
const char *kCrossHelpers = R"VGPU(
template <typename T>
__device__ __forceinline__ void vg_fill_run(T* out, vg_u64 g0, unsigned cnt, vg_u64 cap, T value) {
  constexpr unsigned V = 16 / sizeof(T);
  const vg_u64 end = g0 + cnt < cap ? g0 + cnt : cap;
  if (g0 >= end) return;
  const unsigned sh = (unsigned)(((vg_u64)(out + g0) & 15) / sizeof(T));
  T* vb = out + g0 - sh;
  const unsigned pend = (unsigned)(end - g0) + sh;
  T rep[V];
#pragma unroll
  for (unsigned q = 0; q < V; q++) rep[q] = value;
  uint4 w; __builtin_memcpy(&w, rep, 16);
  for (unsigned k = threadIdx.x; k * V < pend; k += blockDim.x) {
    const unsigned p0 = k * V;
    if (p0 >= sh && p0 + V <= pend) {
      vg_store16((void*)(vb + p0), w);
    } else {
      for (unsigned q = 0; q < V; q++) { const unsigned p = p0 + q; if (p >= sh && p < pend) vb[p] = value; }
    }
  }
}
// stage[p] for p in [sh, sh + cnt) is output element g0 - sh + p, where sh = stage_shift(out, g0)
template <typename T>
__device__ __forceinline__ unsigned vg_stage_shift(const T* out, vg_u64 g0) {
  return (unsigned)(((vg_u64)(out + g0) & 15) / sizeof(T));
}
template <typename T>
__device__ __forceinline__ void vg_copy_run(T* out, vg_u64 g0, unsigned cnt, vg_u64 cap, const T* stage, unsigned sh) {
  constexpr unsigned V = 16 / sizeof(T);
  const vg_u64 end = g0 + cnt < cap ? g0 + cnt : cap;
  if (g0 >= end) return;
  T* vb = out + g0 - sh;
  const unsigned pend = (unsigned)(end - g0) + sh;
  for (unsigned k = threadIdx.x; k * V < pend; k += blockDim.x) {
    const unsigned p0 = k * V;
    if (p0 >= sh && p0 + V <= pend) {
      uint4 w; __builtin_memcpy(&w, __builtin_assume_aligned((const void*)(stage + p0), 16), 16);
      vg_store16((void*)(vb + p0), w);
    } else {
      for (unsigned q = 0; q < V; q++) { const unsigned p = p0 + q; if (p >= sh && p < pend) vb[p] = stage[p]; }
    }
  }
}
)VGPU";

bool LeftOnly(const FusedSpec &s, const FusedExpr &e) {
	for (int slot : e.slots) {
		if (s.cols[slot].side != 0) {
			return false;
		}
	}
	return true;
}

std::string InputRef(const FusedSpec &s, int slot) {
	return s.cols[slot].side == 0 ? "&__l" + std::to_string(slot) : "&__v" + std::to_string(slot) + "[__r][__j]";
}

//! The expression body in a scope of its own: inputs[k] point at the row's values, `out`/`idx` at the target.
void EmitBody(std::ostringstream &o, const FusedSpec &s, const FusedExpr &e, const std::string &inputs_of_slot_kind,
              const std::string &out_ptr, const std::string &idx, const std::string &indent) {
	o << indent << "{\n";
	const size_t k = e.slots.size();
	o << indent << "  const void* inputs[" << (k ? k : 1) << "] = {";
	for (size_t i = 0; i < k; i++) {
		o << (i ? ", " : "") << "(const void*)"
		  << (inputs_of_slot_kind == "left" ? "&__l" + std::to_string(e.slots[i]) : InputRef(s, e.slots[i]));
	}
	if (k == 0) {
		o << "0";
	}
	o << "}; (void)inputs;\n";
	o << indent << "  const vg_u64 row = 0; (void)row;\n";
	o << indent << "  const vg_u64 idx = " << idx << "; (void)idx;\n";
	o << indent << "  void* out = " << out_ptr << "; (void)out;\n";
	o << indent << "  " << e.body << "\n";
	o << indent << "}\n";
}

//! AND of the validity of an expression's inputs ("" when none can be NULL).
std::string ValidOf(const FusedSpec &s, const FusedExpr &e) {
	std::string v;
	for (int slot : e.slots) {
		if (!s.cols[slot].has_valid) {
			continue;
		}
		v += v.empty() ? "" : " && ";
		v += s.cols[slot].side == 0 ? "__lm" + std::to_string(slot) : "__m" + std::to_string(slot) + "[__r][__j]";
	}
	return v;
}

} // namespace

std::string CrossAlignedKernelName(const FusedSpec &s, int tile_pairs, int phase, int rows_per_block) {
	(void)s;
	return "vectis_cross_filter_t" + std::to_string(tile_pairs) + "x" + std::to_string(rows_per_block) +
	       (phase == 1 ? "_count" : "_write");
}

std::string GenerateCrossAlignedSource(const FusedSpec &s, int tile_pairs, int phase, int rows_per_block) {
	if (!s.cross || s.predicates.empty() || s.emit_selection || !s.range_slots.empty() || s.selection_in ||
	    s.ranges) {
		throw std::runtime_error("cross kernel: needs a filtered cross product without selection or range outputs");
	}
	if (tile_pairs % (kCrossBlock * kCrossVec) != 0 || (phase != 1 && phase != 2) || rows_per_block < 1) {
		throw std::runtime_error("cross kernel: bad tile, phase or rows per block");
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
	for (auto &c : s.cols) {
		if (c.segment_id) {
			throw std::runtime_error("cross kernel: no segment-id columns");
		}
	}
	const int rounds = tile_pairs / (kCrossBlock * kCrossVec);
	const int wave_span = tile_pairs / (kCrossBlock / 64);
	std::ostringstream o;

	// The right side's tile-long chunk starting at pair __rj0 into register tiles (4 consecutive rows per lane
	// and round), with 16-byte loads where the chunk is aligned.
	auto emit_right_load = [&](const std::string &in) {
		o << "#pragma unroll\n" << in << "for (int __r = 0; __r < " << rounds << "; __r++) {\n";
		o << in << "  const vg_u64 __rj = __rj0 + (unsigned)wave * " << wave_span << "u + (unsigned)__r * "
		  << (64 * kCrossVec) << "u + (unsigned)lane * " << kCrossVec << "u;\n";
		for (size_t c = 0; c < s.cols.size(); c++) {
			if (s.cols[c].side == 0) {
				continue;
			}
			const char *t = FusedCType(s.cols[c].vtype);
			const std::string cn = "__c" + std::to_string(c), vn = "__v" + std::to_string(c) + "[__r]";
			o << in << "  if ((((vg_u64)(" << cn << " + __rj)) & (" << kCrossVec << " * sizeof(" << t
			  << ") - 1)) == 0) {\n";
			o << in << "    __builtin_memcpy(&" << vn << "[0], __builtin_assume_aligned((const void*)(" << cn
			  << " + __rj), " << kCrossVec << " * sizeof(" << t << ")), " << kCrossVec << " * sizeof(" << t << "));\n";
			o << in << "  } else {\n";
			o << "#pragma unroll\n" << in << "    for (int __j = 0; __j < " << kCrossVec << "; __j++) " << vn
			  << "[__j] = " << cn << "[__rj + __j];\n";
			o << in << "  }\n";
			if (s.cols[c].has_valid) {
				o << "#pragma unroll\n" << in << "  for (int __j = 0; __j < " << kCrossVec << "; __j++) __m" << c
				  << "[__r][__j] = a.valid[" << c << "] == 0 || a.valid[" << c << "][__rj + __j] != 0;\n";
			}
		}
		o << in << "}\n";
	};
	// The left columns' values of left row __li.
	auto emit_left_load = [&](const std::string &in) {
		for (size_t c = 0; c < s.cols.size(); c++) {
			if (s.cols[c].side != 0) {
				continue;
			}
			const char *t = FusedCType(s.cols[c].vtype);
			o << in << "const " << t << " __l" << c << " = __c" << c << "[__li];\n";
			if (s.cols[c].has_valid) {
				o << in << "const bool __lm" << c << " = a.valid[" << c << "] == 0 || a.valid[" << c << "][__li] != 0;\n";
			}
		}
	};
	// __ok = every predicate holds (and its inputs are valid) for pair (__li, right row of register [__r][__j]).
	auto emit_predicates = [&](const std::string &in) {
		o << in << "bool __ok = true;\n";
		for (size_t p = 0; p < s.predicates.size(); p++) {
			const auto &e = s.predicates[p];
			o << in << "bool __p" << p << " = false;\n";
			EmitBody(o, s, e, "row", "(void*)&__p" + std::to_string(p), "0", in);
			const std::string v = ValidOf(s, e);
			o << in << "__ok = __ok && __p" << p << (v.empty() ? "" : " && (" + v + ")") << ";\n";
		}
	};
	// __wc = this wave's kept pairs of left row __li against the chunk in registers (ballots, no barrier).
	auto emit_count = [&](const std::string &in) {
		o << in << "unsigned int __wc = 0;\n";
		o << "#pragma unroll\n" << in << "for (int __r = 0; __r < " << rounds << "; __r++) {\n";
		o << "#pragma unroll\n" << in << "  for (int __j = 0; __j < " << kCrossVec << "; __j++) {\n";
		emit_predicates(in + "    ");
		o << in << "    __wc += (unsigned int)__popcll(__ballot(__ok));\n";
		o << in << "  }\n";
		o << in << "}\n";
	};
	// The kept pairs of left row __li against the chunk in registers, written in order from output row __gbase:
	// left-only outputs as one fill, the others compacted through the LDS stage. Ends with a barrier.
	auto emit_write_row = [&](const std::string &in) {
		o << in << "unsigned int __keep[" << rounds << "];\n";
		o << "#pragma unroll\n" << in << "for (int __r = 0; __r < " << rounds << "; __r++) {\n";
		o << in << "  unsigned int __kb = 0;\n";
		o << "#pragma unroll\n" << in << "  for (int __j = 0; __j < " << kCrossVec << "; __j++) {\n";
		emit_predicates(in + "    ");
		o << in << "    if (__ok) __kb |= 1u << __j;\n";
		o << in << "  }\n";
		o << in << "  __keep[__r] = __kb;\n";
		o << in << "}\n";
		// in-order positions: rounds within the wave's span, waves in order
		o << in << "unsigned int __lpos0[" << rounds << "];\n";
		o << in << "unsigned int __wrun = 0;\n";
		o << "#pragma unroll\n" << in << "for (int __r = 0; __r < " << rounds << "; __r++) {\n";
		o << in << "  unsigned int __below, __tot;\n";
		o << in << "  vg_rank4(__keep[__r], __below, __tot);\n";
		o << in << "  __lpos0[__r] = __wrun + __below;\n";
		o << in << "  __wrun += __tot;\n";
		o << in << "}\n";
		o << in << "if (lane == 0) s_wtot[wave] = __wrun;\n";
		o << in << "__syncthreads();\n";
		o << in << "unsigned int __wexcl = 0, __ttot = 0;\n";
		o << in << "for (int w = 0; w < 4; w++) { if (w < wave) __wexcl += s_wtot[w]; __ttot += s_wtot[w]; }\n";
		o << "#pragma unroll\n" << in << "for (int __r = 0; __r < " << rounds << "; __r++) __lpos0[__r] += __wexcl;\n";
		for (size_t e = 0; e < s.projections.size(); e++) {
			const auto &ex = s.projections[e];
			const char *ot = FusedCType(ex.out_type);
			const std::string out = "((" + std::string(ot) + "*)a.outs[" + std::to_string(e) + "])";
			const std::string v = ValidOf(s, ex);
			if (LeftOnly(s, ex)) {
				// one value for every kept pair of the tile
				o << in << "{\n";
				o << in << "  " << ot << " __u;\n";
				EmitBody(o, s, ex, "left", "(void*)&__u", "0", in + "  ");
				o << in << "  vg_fill_run<" << ot << ">(" << out << ", __gbase, __ttot, a.cap, __u);\n";
				if (!v.empty()) {
					o << in << "  if (a.out_valid[" << e << "] != 0) vg_fill_run<unsigned char>(a.out_valid[" << e
					  << "], __gbase, __ttot, a.cap, (unsigned char)((" << v << ") ? 1 : 0));\n";
				}
				o << in << "}\n";
				continue;
			}
			auto emit_staged = [&](const std::string &type, const std::string &target,
			                       const std::function<void()> &value) {
				o << in << "{\n";
				o << in << "  " << type << "* __st = (" << type << "*)s_stage;\n";
				o << in << "  const unsigned __sh = vg_stage_shift<" << type << ">(" << target << ", __gbase);\n";
				o << "#pragma unroll\n" << in << "  for (int __r = 0; __r < " << rounds << "; __r++) {\n";
				o << in << "    unsigned int __lp = __lpos0[__r] + __sh;\n";
				o << "#pragma unroll\n" << in << "    for (int __j = 0; __j < " << kCrossVec << "; __j++) {\n";
				o << in << "      if ((__keep[__r] >> __j) & 1u) {\n";
				o << in << "        const unsigned __o = __lp++;\n";
				value();
				o << in << "      }\n";
				o << in << "    }\n";
				o << in << "  }\n";
				o << in << "  __syncthreads();\n";
				o << in << "  vg_copy_run<" << type << ">(" << target << ", __gbase, __ttot, a.cap, __st, __sh);\n";
				o << in << "  __syncthreads();\n";
				o << in << "}\n";
			};
			emit_staged(ot, out, [&]() { EmitBody(o, s, ex, "row", "(void*)__st", "__o", in + "        "); });
			if (!v.empty()) {
				o << in << "if (a.out_valid[" << e << "] != 0)\n";
				emit_staged("unsigned char", "a.out_valid[" + std::to_string(e) + "]",
				            [&]() { o << in << "        __st[__o] = (" << v << ") ? 1 : 0;\n"; });
			}
		}
		o << in << "__syncthreads(); // s_wtot and the stage are reused by the next left row\n";
	};

	o << FusedPreamble() << kStore16 << kCrossHelpers;
	o << "extern \"C\" __global__ void __launch_bounds__(" << kCrossBlock << ") "
	  << CrossAlignedKernelName(s, tile_pairs, phase, rows_per_block) << "(VArgs a) {\n";
	o << "  const int tid = threadIdx.x; const int lane = tid & 63; const int wave = tid >> 6;\n";
	o << "  const vg_u64 __chunks = a.m / " << tile_pairs << "ull;\n";
	// block = (right chunk, group of rows_per_block left rows); tile t = left row x chunks + right chunk
	o << "  const vg_u64 __chunk = blockIdx.x % __chunks;\n";
	o << "  const vg_u64 __li0 = (blockIdx.x / __chunks) * " << rows_per_block << "ull;\n";
	o << "  const vg_u64 __nl = a.n / a.m;\n";
	o << "  const vg_u64 __rj0 = __chunk * " << tile_pairs << "ull;\n";
	o << "  __shared__ unsigned int s_wtot[4];\n";
	// right columns: register tiles of 4 consecutive rows per round
	for (size_t c = 0; c < s.cols.size(); c++) {
		const char *t = FusedCType(s.cols[c].vtype);
		o << "  const " << t << "* __c" << c << " = (const " << t << "*)a.cols[" << c << "];\n";
		if (s.cols[c].side != 0) {
			o << "  " << t << " __v" << c << "[" << rounds << "][" << kCrossVec << "];\n";
			if (s.cols[c].has_valid) {
				o << "  bool __m" << c << "[" << rounds << "][" << kCrossVec << "];\n";
			}
		}
	}
	// loaded once for all the block's left rows
	emit_right_load("  ");
	if (phase == 2) {
		int stage_width = 1;
		for (auto &e : s.projections) {
			if (!LeftOnly(s, e)) {
				stage_width = std::max(stage_width, TypeWidth(e.out_type));
			}
		}
		o << "  __shared__ __attribute__((aligned(16))) vg_u64 s_stage["
		  << ((uint64_t)tile_pairs * stage_width + 16 + 7) / 8 << "];\n";
	}
	if (phase == 1) {
		// count pass: per left row, each wave adds popc(ballot(kept)) over its pairs (scalar work, no barrier);
		// one barrier at the end, then one thread per left row sums the four waves
		o << "  __shared__ unsigned int s_cnt[4][" << rows_per_block << "];\n";
		o << "#pragma unroll 1\n  for (int __k = 0; __k < " << rows_per_block << "; __k++) {\n";
		o << "    const vg_u64 __li = __li0 + __k;\n";
		o << "    if (__li >= __nl) break;\n";
		emit_left_load("    ");
		emit_count("    ");
		o << "    if (lane == 0) s_cnt[wave][__k] = __wc;\n";
		o << "  }\n";
		o << "  __syncthreads();\n";
		o << "  if (tid < " << rows_per_block << " && __li0 + tid < __nl) {\n";
		o << "    a.state[(__li0 + tid) * __chunks + __chunk] = (vg_u64)(s_cnt[0][tid] + s_cnt[1][tid] + s_cnt[2][tid] + "
		     "s_cnt[3][tid]);\n";
		o << "  }\n";
		o << "}\n";
		return o.str();
	}
	o << "#pragma unroll 1\n  for (int __k = 0; __k < " << rows_per_block << "; __k++) {\n";
	o << "  const vg_u64 __li = __li0 + __k;\n";
	o << "  if (__li >= __nl) break;\n";
	o << "  const vg_u64 tile = __li * __chunks + __chunk;\n";
	emit_left_load("  ");
	o << "  const vg_u64 __gbase = a.state[tile];\n";
	emit_write_row("  ");
	o << "  }\n";
	o << "}\n";
	return o.str();
}

} // namespace vgpu
