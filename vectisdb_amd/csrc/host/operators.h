// operators.h -- device-resident operator implementations shared by the plan executor (GpuEngine) and the
// extern "C" boundary. Inputs and outputs are device columns; nothing here touches host data.
#pragma once

#include "../hip/kernels.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
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
uint64_t PlanGroupBy(const std::vector<DevCol> &keys, uint64_t rows, KeyPack &pack, Scratch &scratch,
                     hipStream_t stream, const uint32_t **hist8 = nullptr);

//! Sort-based GROUP BY. Key outputs (one per key column, key type) and aggregate outputs (out_type) are
//! allocated through `out_alloc` with `capacity` rows each. Returns the group count when `read_groups`
//! (synchronizes), else writes it to *groups_device and returns 0.
uint64_t GroupByDevice(const std::vector<DevCol> &keys, const std::vector<DevCol> &values,
                       const std::vector<AggReq> &aggs, uint64_t rows, const KeyPack &pack, uint64_t capacity,
                       const OutAlloc &out_alloc, std::vector<void *> &key_out, std::vector<void *> &agg_out,
                       const SortTuning &tuning, Scratch &scratch, hipStream_t stream, bool read_groups,
                       uint64_t *groups_device = nullptr);

//! Autotuner: radix digit width per key width / row count, from measured sweeps (see tuning.cpp).
SortTuning TuneSort(int total_bits, uint64_t rows);

} // namespace vgpu
