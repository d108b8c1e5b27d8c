#pragma once
// gpu_column_cache.hpp -- device-resident column cache (drop-in for cuda_engine/include/gpu_column_cache.hpp).
//
// Columns live in HBM as a sequence of chunks (one per scan batch). A scan stages chunks into an open slot
// (BeginStaging / StageChunk) and publishes them only when CommitStaging confirms that every column holds
// exactly the scanned row count; anything else is discarded. Committed columns are evicted least recently used
// first under a byte budget sized for a 288 GB MI355X: three quarters of the free HBM at first use, always
// leaving 24 GiB free for operator intermediates (about 210 GB of tables; VECTOR_GPU_CACHE_FRACTION /
// VECTOR_GPU_CACHE_MB override). Device pointers handed out are borrows, valid until the next call that can
// evict.

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "gpu_engine.hpp"

namespace vector_gpu {

struct GpuColumnChunk {
	void *device_data = nullptr;     // device memory; owned by the cache once StageChunk returns true
	void *device_validity = nullptr; // dense byte validity (1 = valid), nullptr when the chunk has no NULLs
	uint64_t row_count = 0;
	uint64_t data_bytes = 0;
	uint64_t validity_bytes = 0;
	GpuValueType type = GpuValueType::FLOAT64;

	uint64_t TotalBytes() const {
		return data_bytes + validity_bytes;
	}
};

class GpuColumnCache {
public:
	static GpuColumnCache &Instance();

	GpuColumnCache(const GpuColumnCache &) = delete;
	GpuColumnCache &operator=(const GpuColumnCache &) = delete;

	//! Sizes the budget once: min(fraction_of_vram x free HBM, free HBM - 24 GiB) (env overrides).
	void EnsureInitialized(double fraction_of_vram = 0.75);

	//! Opens (or re-opens, discarding) the staging slot of `table`.
	void BeginStaging(const GpuTableRef &table);
	//! Appends a chunk to the column's staged sequence. True = the cache now owns the buffers.
	bool StageChunk(const GpuTableRef &table, const std::string &column_name, const GpuColumnChunk &chunk);
	//! Publishes every staged column whose row count equals `scanned_rows`; frees the rest.
	void CommitStaging(const GpuTableRef &table, uint64_t scanned_rows);
	void AbortStaging(const GpuTableRef &table);

	//! Per-chunk row counts when every named column is resident with identical chunking; empty otherwise.
	std::vector<uint64_t> ReplayableChunkRows(const GpuTableRef &table, const std::vector<std::string> &column_names);
	bool TryGetChunk(const GpuTableRef &table, const std::string &column_name, uint64_t chunk_index,
	                 GpuColumnChunk &out_chunk);

	void InvalidateTable(const GpuTableRef &table);
	void InvalidateAll();

	uint64_t CurrentBytes() const;
	uint64_t TableGeneration(const GpuTableRef &table) const;

private:
	GpuColumnCache();
	~GpuColumnCache();

	// A committed column: its chunks, and its place in the recency index.
	struct Resident {
		std::vector<GpuColumnChunk> chunks;
		uint64_t rows = 0;
		uint64_t bytes = 0;
		uint64_t last_use = 0;
	};
	// A column being staged by an open scan.
	struct Pending {
		std::vector<GpuColumnChunk> chunks;
		uint64_t rows = 0;
		uint64_t bytes = 0;
		GpuValueType type = GpuValueType::FLOAT64;
		bool typed = false;
		bool dead = false; // a chunk was refused: the column cannot be committed by this scan
	};
	struct Table {
		uint64_t generation = 0;
		bool staging = false;
		std::map<std::string, Resident> resident;
		std::map<std::string, Pending> pending;
	};
	using Key = std::string; // catalog, schema and table joined by NUL bytes

	static Key KeyOf(const GpuTableRef &table);
	void EnsureInitializedLocked(double fraction_of_vram);
	bool MakeRoomLocked(uint64_t need);
	void DropResidentLocked(Table &table, std::map<std::string, Resident>::iterator it);
	void DropPendingLocked(Table &table);
	void UseLocked(const Key &key, const std::string &column, Resident &r);

	mutable std::mutex mutex_;
	bool initialized_ = false;
	uint64_t budget_bytes_ = 0;
	uint64_t resident_bytes_ = 0;
	uint64_t pending_bytes_ = 0;
	uint64_t use_clock_ = 0;
	std::map<Key, Table> tables_;
	// least recently used first: use tick -> (table key, column)
	std::map<uint64_t, std::pair<Key, std::string>> recency_;
};

} // namespace vector_gpu
