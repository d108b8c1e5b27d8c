// operators.h -- device-resident operator implementations shared by the plan executor (GpuEngine) and the
// extern "C" boundary. Inputs and outputs are device columns; nothing here touches host data.
#pragma once

#include "../engine/fused.h"
#include "../hip/kernels.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <memory>
#include <vector>

namespace vgpu {

struct DevCol {
	const void *data = nullptr;
	int type = VT_INT64;
	uint64_t rows = 0;
};

struct AggReq {
	int kind = AGG_COUNT_STAR; // AggKind (== GpuAggregateKind numbering)
	int input = -1;            // index into the value columns; -1 for COUNT_STAR
	int out_type = VT_INT64;
};

using OutAlloc = std::function<void *(size_t bytes)>;

//! Validates one aggregate the way the engine's GROUP BY accepts it; throws std::runtime_error with a
//! message naming the offending rule. Extends the reference rules with integer SUM -> HUGEINT (DuckDB's
//! own result type) and floating MIN/MAX with DuckDB's NaN ordering.
void ValidateAggregateReq(const AggReq &agg, const std::vector<DevCol> &values);

//! Upper bound on the number of groups (min of rows and the product of key ranges); also returns the key
//! packing. Synchronizes (reads key ranges back). Throws when the keys need more than 64 packed bits.
//! For a single key column the range pass also builds the byte-digit histograms (*hist8, from `scratch`),
//! which the sort then reuses instead of reading the keys again.
//! For one key column, *span_out (when given) receives max - min + 1 of its values (0 when that is 2^64).
uint64_t PlanGroupBy(const std::vector<DevCol> &keys, uint64_t rows, KeyPack &pack, Scratch &scratch,
                     hipStream_t stream, const uint32_t **hist8 = nullptr, uint64_t *span_out = nullptr);
//! The same plan from key ranges already known (order-preserving min/max per key column, as the key-range
//! pass returns them); reads nothing from the device.
uint64_t PlanGroupByRanges(const std::vector<DevCol> &keys, uint64_t rows, const uint64_t *mins,
                           const uint64_t *maxs, KeyPack &pack);

//! Sort-based GROUP BY. Key outputs (one per key column, key type) and aggregate outputs (out_type) are
//! allocated through `out_alloc` with `capacity` rows each. Returns the group count when `read_groups`
//! (synchronizes), else writes it to *groups_device and returns 0. With rows_dev the row count is read on
//! the device (an earlier kernel on `stream` wrote it; `rows` and the columns' row counts are its bound),
//! so a call with read_groups = false makes no host read and no synchronizing call.
//!
//! narrow_payload (one key column, read_groups, host row count): INT64 aggregate inputs are carried through
//! the sort as their low 32 bits, which moves 8 bytes less per row in every pass and in the reduction. The
//! first pass checks every value against the INT32 range; if one is outside it, the GROUP BY runs again
//! without narrowing into the same outputs, and the column (address, rows) is remembered so later calls on
//! it do not narrow.
uint64_t GroupByDevice(const std::vector<DevCol> &keys, const std::vector<DevCol> &values,
                       const std::vector<AggReq> &aggs, uint64_t rows, const KeyPack &pack, uint64_t capacity,
                       const OutAlloc &out_alloc, std::vector<void *> &key_out, std::vector<void *> &agg_out,
                       const SortTuning &tuning, Scratch &scratch, hipStream_t stream, bool read_groups,
                       uint64_t *groups_device = nullptr, const uint64_t *rows_dev = nullptr,
                       bool narrow_payload = false);

//! The GROUP BY's last radix pass fused with the reduction (RunReduceLastPass): 0 never, 1 when the partial
//! states it may write are fewer bytes than the rows a plain last pass scatters (the default, or the value of
//! VGPU_RUN_REDUCE at first use), 2 whenever the aggregate mix is supported.
void SetRunReduceMode(int mode);
int RunReduceMode();

// ------------------------------------------------------------------ join

struct JoinAggReq {
	int kind = AGG_COUNT_STAR; // AggKind
	int side = -1;             // 0 = A (children[0]), 1 = B (children[1]); -1 for COUNT(*)
	int input = -1;            // index into that side's value columns
	int out_type = VT_INT64;
};

//! One aggregate over the join, as its raw two-word state: COUNT -> (lo, hi) uint128; SUM int -> int128;
//! SUM float -> double-double (hi word = bits of the high double); MIN/MAX -> lo.
struct JoinAggValue {
	uint64_t lo = 0;
	uint64_t hi = 0;
	int kind = 0; // JoinAggKind
};

//! Aggregate-only INNER equi-join (global aggregate over the join, nothing materialized). Synchronizes.
std::vector<JoinAggValue> JoinAggregateDevice(const DevCol &akey, const std::vector<DevCol> &avals, const DevCol &bkey,
                                              const std::vector<DevCol> &bvals, const std::vector<JoinAggReq> &aggs,
                                              Scratch &scratch, hipStream_t stream);

//! Writes one aggregate-over-join result as its output type (TypeWidth(out_type) bytes at `dst`): COUNT as
//! INT64, integer SUM as HUGEINT / INT64 (low word) / DOUBLE, floating SUM as DOUBLE (double-double rounded
//! once), AVG as SUM / `pairs`, MIN/MAX as the input's own 8-byte value.
void FinishJoinAggregate(const JoinAggValue &v, int agg_kind, int out_type, uint64_t pairs, void *dst);

//! Materialized INNER equi-join: returns the number of output rows and, in *left_rows / *right_rows (from
//! `scratch`), the source row of each side per output row, ordered by key, then left row, then right row.
//! Synchronizes. Throws past 2^32 output rows.
uint64_t JoinMaterializeDevice(const DevCol &akey, const DevCol &bkey, uint32_t **left_rows, uint32_t **right_rows,
                               Scratch &scratch, hipStream_t stream);

// ------------------------------------------------------------------ batched form

//! Many filter + GROUP BY queries over the row groups of one table (segment g = rows
//! [offsets[g], offsets[g+1])), run as one pipeline. Predicate slots index `cols`; aggregate inputs index
//! `value_cols`. Outputs (caller-owned, `capacity` rows each): the segment of each group (INT32), the key,
//! and one column per aggregate; groups ordered by (segment, key).
struct BatchedGroupByArgs {
	std::vector<DevCol> cols;
	uint64_t rows = 0;
	const uint64_t *seg_offsets_dev = nullptr;
	uint32_t nseg = 0;
	//! Row ranges instead of contiguous row groups (both host arrays of nseg entries, or both null): query g
	//! reads table rows [range_rows[g], range_rows[g] + range_lens[g]); ranges may overlap or repeat.
	const uint64_t *range_rows_host = nullptr;
	const uint64_t *range_lens_host = nullptr;
	std::vector<FusedExpr> preds;
	int key_col = -1;
	std::vector<int> value_cols;
	std::vector<AggReq> aggs;
	void *seg_out = nullptr;
	void *key_out = nullptr;
	std::vector<void *> agg_out;
	uint64_t capacity = 0;
	uint64_t *seg_group_offsets_host = nullptr; // nseg + 1 entries, optional
	bool fast_math = false;
};

//! Returns the total group count. Synchronizes.
uint64_t BatchedFilterGroupByDevice(const BatchedGroupByArgs &args, Scratch &scratch, hipStream_t stream);

//! The same queries run one at a time (filter kernel, GROUP BY, each with its own size read-backs), writing
//! the same output layout: the plain per-query path the batched form is measured against.
//! `seg_offsets_host` holds the nseg + 1 row offsets on the host.
uint64_t PerQueryFilterGroupByDevice(const BatchedGroupByArgs &args, const uint64_t *seg_offsets_host,
                                     Scratch &scratch, hipStream_t stream);

//! Row ranges of one launch on the device: range g covers launch rows [seg[g], seg[g] + lens[g]) read from
//! table rows rows[g] + (v - seg[g]). Launch starts are padded so seg[g] = rows[g] (mod 4): a lane's four
//! rows stay 4-aligned in the table, which keeps the fused kernel's 128-bit loads.
struct DeviceRanges {
	const uint64_t *seg = nullptr;  // nseg + 1
	const uint64_t *rows = nullptr; // nseg
	const uint64_t *lens = nullptr; // nseg
	uint64_t nvirt = 0;             // launch rows, seg[nseg]
	uint32_t nseg = 0;
};
//! Checks every range against `table_rows` and uploads the launch layout (synchronizes the stream).
DeviceRanges UploadRanges(const uint64_t *rows, const uint64_t *lens, uint32_t nseg, uint64_t table_rows,
                          Scratch &scratch, hipStream_t stream);

//! Many filter/projection queries over row ranges of one table in one generated kernel. Outputs (from
//! `out_scratch`, one per projection) hold every query's kept rows, query after query in range order;
//! query g's rows are [kept_offsets[g], kept_offsets[g+1]). Returns the total kept. Synchronizes.
uint64_t BatchedFilterProjectDevice(const std::vector<DevCol> &cols, const DeviceRanges &ranges,
                                    const std::vector<FusedExpr> &preds, const std::vector<FusedExpr> &projs,
                                    bool fast_math, std::vector<void *> &outs, std::vector<uint64_t> &kept_offsets,
                                    Scratch &out_scratch, Scratch &scratch, hipStream_t stream);

//! Build-side index for aggregate-only joins: the distinct build keys ascending (key type), each with its
//! row count (uint64) and the int128 sum of its payload (lo, hi words). Device memory from DevicePool,
//! owned by the index.
struct JoinIndex {
	int key_type = VT_INT64;
	void *keys = nullptr;
	uint64_t *count = nullptr;
	uint64_t *sum = nullptr;
	uint64_t distinct = 0;
	uint64_t build_rows = 0;
	//! Direct-address table (BuildDenseIndex) when the keys are dense enough, else null.
	uint64_t *dense = nullptr;
	uint64_t dense_min = 0;  // ordered as uint64 of the sign-extended key
	uint64_t dense_span = 0;
	JoinIndex() = default;
	JoinIndex(const JoinIndex &) = delete;
	JoinIndex &operator=(const JoinIndex &) = delete;
	~JoinIndex();
};
//! Reduces the build side (integer key, integer payload, no NULLs) with the GROUP BY operator; when the
//! distinct keys span at most twice their number (plus 1M) and every count and sum packs into 64 bits, also
//! builds the direct-address table.
//! A payload with data == nullptr builds a count-only index (every sum 0). With dense_only, an index whose
//! keys span more than 2 x rows + 2^20 values (so no direct-address table can be built) is returned
//! without being built (distinct 0, dense null), skipping the sort.
std::unique_ptr<JoinIndex> BuildJoinIndex(const DevCol &key, const DevCol &payload, hipStream_t stream,
                                          bool dense_only = false);

//! SUM(payload) and COUNT(*) of key = index key for every probe range, as JoinAggValue pairs
//! (result[2g] = SUM as int128, result[2g+1] = COUNT). Synchronizes.
std::vector<JoinAggValue> BatchedJoinAggDevice(const DevCol &probe_key, const DeviceRanges &ranges,
                                               const JoinIndex &index, Scratch &scratch, hipStream_t stream);

//! Autotuner: radix digit width per key width / row count, from measured sweeps (see tuning.cpp).
SortTuning TuneSort(int total_bits, uint64_t rows);

} // namespace vgpu
