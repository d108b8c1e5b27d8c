// fused.h -- generated fused filter/projection kernels (hipRTC, gfx950).
//
// One kernel evaluates every predicate of a FILTER (AND-ed; a NULL predicate drops the row) and every
// projection above it, compacts the surviving rows in order (wave64 ballots + per-lane prefix, global
// offset by decoupled look-back) and writes dense outputs -- one read of each input column and one write
// of each output, instead of the reference's mask kernel per predicate, AND kernel, copy_if, narrowing copy
// and one kernel per projection reading through a selection vector
// (cuda_engine/src/gpu_executor.cu:645-746).
//
// Expression bodies are the extension's text unchanged (they read ((const T*)inputs[k])[row] and write
// ((U*)out)[idx]); the generator gives each body its own inputs[] of pointers into per-row registers, so
// the same text runs over 128-bit-loaded register tiles, over a gathered selection, or over the two sides
// of a cross product.
#pragma once

#include "rtc.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace vgpu {

struct FusedColumn {
	int vtype = 1;          // VType of the column
	bool has_valid = false; // dense byte validity supplied at launch
	int side = 0;           // cross-product mode: 0 = left (broadcast), 1 = right
	//! Not read from memory: the row's segment index g (seg_offsets[g] <= row < seg_offsets[g+1]), as INT32.
	bool segment_id = false;
};

struct FusedExpr {
	std::string body;
	std::vector<int> slots; // inputs[k] = column slots[k]
	int out_type = 4;
};

struct FusedSpec {
	std::vector<FusedColumn> cols;
	std::vector<FusedExpr> predicates;
	std::vector<FusedExpr> projections;
	bool selection_in = false;   // rows are read through a uint32 selection vector
	bool emit_selection = false; // write kept row ids (uint32, base-row numbering)
	bool cross = false;          // rows are (left i, right j) pairs of a left-major cross product
	//! Rows are row ranges of the table: segment g covers launch rows [seg[g], seg[g] + seg_len[g]), read
	//! from table rows seg_row[g] + (row - seg[g]); launch rows from seg[g] + seg_len[g] up to seg[g+1] are
	//! padding (never kept). Needs at least one predicate (the outputs are compacted).
	bool ranges = false;
	//! With `ranges`: the launch's row and tile counts are read from device memory (FusedLaunch::dims_device,
	//! written by an earlier kernel), and a fixed grid of resident blocks sweeps the tiles, so the launch
	//! needs nothing from the host that depends on the data (it can be captured into a graph). The caller
	//! zeroes the look-back words and the ticket (state[0 .. ntiles]) and the count before the launch.
	bool persistent = false;
	bool fast_math = false;
	//! Non-null integer columns (a segment-id column included) whose minimum and maximum over the kept rows
	//! the launch writes to `range_device`, as order-preserving unsigned keys (signed values sign-extended
	//! to 64 bits with the top bit flipped).
	std::vector<int> range_slots;
};

struct FusedLaunch {
	const void *const *cols = nullptr;        // host array of device column pointers (cols.size())
	const uint8_t *const *valid = nullptr;    // host array (cols.size()) of dense validity or nullptr
	void *const *outs = nullptr;              // host array of projection outputs
	uint8_t *const *out_valid = nullptr;      // host array (projections) or nullptr entries
	uint64_t n = 0;                           // rows (or selection length, or left rows x right rows)
	uint64_t right_rows = 0;                  // cross mode: m
	const uint32_t *sel = nullptr;            // selection_in
	uint32_t *sel_out = nullptr;              // emit_selection
	uint64_t *count_device = nullptr;         // kept rows (device); required when predicates exist
	uint64_t capacity = 0;                    // output rows available (0 = n); writes past it are dropped
	const uint64_t *seg_offsets = nullptr;    // device, nseg + 1 ascending row offsets (segment-id columns)
	uint64_t nseg = 0;
	const uint64_t *seg_rows = nullptr;       // ranges mode: device, nseg table rows of each segment's start
	const uint64_t *seg_lens = nullptr;       // ranges mode: device, nseg segment lengths
	//! range_slots: device, 2 words per slot, written as {min, ~max} (both kept by atomic minimum; the
	//! launch sets them to ~0 first, so a launch that keeps no row leaves {~0, ~0}).
	uint64_t *range_device = nullptr;
	//! Persistent form: device {launch rows, tiles of tile_rows() rows}; `n` is then the bound the state
	//! buffer is sized for.
	const uint64_t *dims_device = nullptr;
};

class FusedKernel {
public:
	~FusedKernel();
	//! Launches on `stream`. Pointer tables are passed by value in the kernel arguments (no device-side
	//! scratch, so concurrent launches on different streams never share state).
	void Launch(const FusedLaunch &args, void *scratch_state, size_t scratch_bytes, hipStream_t stream) const;
	//! Device bytes of look-back state a launch over `n` rows needs.
	static size_t StateBytes(uint64_t n);
	const FusedSpec &spec() const {
		return spec_;
	}
	//! Rows per tile of the default variant (the persistent form always runs it).
	uint64_t tile_rows() const;
	//! Blocks of the persistent form's grid (resident blocks per CU x CUs).
	unsigned persistent_grid() const {
		return persistent_grid_;
	}
	std::string source;

private:
	friend std::shared_ptr<FusedKernel> GetFusedKernel(const FusedSpec &spec);
	FusedSpec spec_;
	RtcKernel *rtc_ = nullptr;
	//! Filtered cross product with the right side a multiple of cross_tile_ pairs (its inputs come from
	//! cache, so evaluating the predicate twice costs less than a tile ticket and a look-back per tile):
	//! rtc_count_ counts each static tile's kept pairs, a scan turns the counts into offsets,
	//! rtc_write_ writes. Other cross products take the single-pass rtc_.
	RtcKernel *rtc_count_ = nullptr;
	RtcKernel *rtc_write_ = nullptr;
	int cross_tile_ = 0;
	int cross_count_rows_ = 0; // left rows per block of the count kernel
	int cross_write_rows_ = 0; // ... and of the write kernel
	//! Variant with larger tiles for large inputs, compiled on its first use (0 rounds: none).
	int large_rounds_ = 0;
	unsigned persistent_grid_ = 0;
	mutable std::mutex large_mutex_;
	mutable RtcKernel *rtc_large_ = nullptr;
};

//! Builds (or fetches from the process cache, keyed by SHA-256 of the generated source) the kernel.
std::shared_ptr<FusedKernel> GetFusedKernel(const FusedSpec &spec);

//! The generated source (exposed for tests and the disk-cache key); rounds_override > 0 sets the tile's
//! row rounds instead of the spec's default.
std::string GenerateFusedSource(const FusedSpec &spec, int rounds_override = 0);
//! The generated kernel's symbol, which names its form and tile so profiles attribute each variant:
//! vectis_{filter_project|project|gather_filter|gather_project|cross_filter|cross_expand|batched_filter}_t<rows>.
std::string FusedKernelName(const FusedSpec &spec, int rounds_override = 0);

//! The helpers compiled into every generated kernel (VArgs, look-back, wave scan), and the C type of a VType.
const char *FusedPreamble();
const char *FusedCType(int vtype);

//! Cross product with a filter whose right side is a whole number of tiles (one left row per tile): the
//! count pass (phase 1) or the write pass (phase 2) over static tiles of `tile_pairs` pairs. Left columns are
//! read once per tile; an output that reads only left columns is the same value for every kept pair of the
//! tile and is written as a fill; the others are compacted through LDS, staged at the output's 16-byte
//! alignment so the copy-out is 16-byte stores. Not for specs with a selection output or range columns.
//! Grid: (m / tile_pairs) x ceil(left rows / rows_per_block) blocks; each block holds one tile-long chunk of the
//! right side in registers and runs it against rows_per_block consecutive left rows.
std::string GenerateCrossAlignedSource(const FusedSpec &spec, int tile_pairs, int phase, int rows_per_block);
std::string CrossAlignedKernelName(const FusedSpec &spec, int tile_pairs, int phase, int rows_per_block);

constexpr int kFusedMaxCols = 16;
constexpr int kFusedMaxOuts = 16;

} // namespace vgpu
