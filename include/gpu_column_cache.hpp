#pragma once
// gpu_column_cache.hpp -- device-resident column cache (drop-in for cuda_engine/include/gpu_column_cache.hpp).
//
// Columns live in HBM as a sequence of chunks (one per scan batch). A scan stages chunks into an open slot
// (BeginStaging / StageChunk) and publishes them only when CommitStaging confirms that every column holds
// exactly the scanned row count; anything else is discarded. Committed columns are evicted LRU under a byte
// budget (default half of free HBM: ~140 GB on one MI355X; VECTOR_GPU_CACHE_FRACTION / VECTOR_GPU_CACHE_MB
// override). Device pointers handed out are borrows, valid until the next call that can evict.

#include <cstdint>
#include <functional>
#include <list>
#include <mutex>
#include <string>
#include <unordered_map>
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

	//! Sizes the budget once: fraction_of_vram x free device memory (env overrides).
	void EnsureInitialized(double fraction_of_vram = 0.5);

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

	struct CachedColumn {
		std::vector<GpuColumnChunk> chunks;
		uint64_t row_count = 0;
		uint64_t total_bytes = 0;
		uint64_t generation = 0;
	};
	struct StagedColumn {
		std::vector<GpuColumnChunk> chunks;
		uint64_t row_count = 0;
		uint64_t total_bytes = 0;
		GpuValueType type = GpuValueType::FLOAT64;
		bool type_known = false;
		bool failed = false;
	};
	struct SlotKey {
		GpuTableRef table;
		std::string column;
		bool operator==(const SlotKey &other) const {
			return table == other.table && column == other.column;
		}
	};
	struct TableRefHash {
		size_t operator()(const GpuTableRef &ref) const {
			std::hash<std::string> h;
			size_t seed = h(ref.catalog);
			seed = seed * 1000003u ^ h(ref.schema);
			seed = seed * 1000003u ^ h(ref.table);
			return seed;
		}
	};
	struct SlotKeyHash {
		size_t operator()(const SlotKey &key) const {
			return TableRefHash {}(key.table) * 31u ^ std::hash<std::string> {}(key.column);
		}
	};

	void EvictLocked(const SlotKey &slot);
	void EvictTableLocked(const GpuTableRef &table);
	void AbortStagingLocked(const GpuTableRef &table);
	void FailStagedColumnLocked(StagedColumn &staged);
	void TouchLocked(const SlotKey &slot);

	mutable std::mutex mutex_;
	bool initialized_ = false;
	uint64_t budget_bytes_ = 0;
	uint64_t committed_bytes_ = 0;
	uint64_t staged_bytes_ = 0;
	std::list<SlotKey> lru_order_;
	std::unordered_map<SlotKey, std::pair<CachedColumn, std::list<SlotKey>::iterator>, SlotKeyHash> entries_;
	std::unordered_map<GpuTableRef, std::unordered_map<std::string, StagedColumn>, TableRefHash> staging_;
	std::unordered_map<GpuTableRef, uint64_t, TableRefHash> generation_;
};

} // namespace vector_gpu
