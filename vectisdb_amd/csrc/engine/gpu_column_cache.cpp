// gpu_column_cache.cpp -- staging/commit protocol and LRU budget of the device-resident column cache
// (behaviour of cuda_engine/src/gpu_column_cache.cpp:61-328; storage is this engine's device pool).
#include "../../../include/gpu_column_cache.hpp"

#include "../host/device_pool.h"

#include <hip/hip_runtime.h>

#include <cstdlib>

namespace vector_gpu {

namespace {
void FreeChunk(GpuColumnChunk &chunk) {
	vgpu::DevicePool::Instance().FreeAny(chunk.device_data);
	vgpu::DevicePool::Instance().FreeAny(chunk.device_validity);
	chunk.device_data = nullptr;
	chunk.device_validity = nullptr;
}
void FreeChunks(std::vector<GpuColumnChunk> &chunks) {
	for (auto &c : chunks) {
		FreeChunk(c);
	}
	chunks.clear();
}
} // namespace

GpuColumnCache &GpuColumnCache::Instance() {
	static GpuColumnCache *cache = new GpuColumnCache();
	return *cache;
}

GpuColumnCache::GpuColumnCache() = default;
GpuColumnCache::~GpuColumnCache() = default;

void GpuColumnCache::EnsureInitialized(double fraction_of_vram) {
	std::lock_guard<std::mutex> guard(mutex_);
	if (initialized_) {
		return;
	}
	initialized_ = true;
	if (const char *env = std::getenv("VECTOR_GPU_CACHE_FRACTION")) {
		char *end = nullptr;
		double v = std::strtod(env, &end);
		if (end != env && v > 0.0 && v <= 1.0) {
			fraction_of_vram = v;
		}
	}
	if (!(fraction_of_vram > 0.0 && fraction_of_vram <= 1.0)) {
		fraction_of_vram = 0.5;
	}
	if (const char *env = std::getenv("VECTOR_GPU_CACHE_MB")) {
		char *end = nullptr;
		unsigned long long mb = std::strtoull(env, &end, 10);
		if (end != env && mb > 0) {
			budget_bytes_ = (uint64_t)mb * 1024ull * 1024ull;
			return;
		}
	}
	size_t free_bytes = 0, total_bytes = 0;
	if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) {
		(void)hipGetLastError();
		budget_bytes_ = 0;
		return;
	}
	budget_bytes_ = (uint64_t)((double)free_bytes * fraction_of_vram);
}

void GpuColumnCache::TouchLocked(const SlotKey &slot) {
	auto it = entries_.find(slot);
	if (it != entries_.end()) {
		lru_order_.splice(lru_order_.begin(), lru_order_, it->second.second);
	}
}

void GpuColumnCache::EvictLocked(const SlotKey &slot) {
	auto it = entries_.find(slot);
	if (it == entries_.end()) {
		return;
	}
	FreeChunks(it->second.first.chunks);
	committed_bytes_ -= it->second.first.total_bytes;
	lru_order_.erase(it->second.second);
	entries_.erase(it);
}

void GpuColumnCache::EvictTableLocked(const GpuTableRef &table) {
	std::vector<SlotKey> victims;
	for (auto &kv : entries_) {
		if (kv.first.table == table) {
			victims.push_back(kv.first);
		}
	}
	for (auto &v : victims) {
		EvictLocked(v);
	}
}

void GpuColumnCache::FailStagedColumnLocked(StagedColumn &staged) {
	FreeChunks(staged.chunks);
	staged_bytes_ -= staged.total_bytes;
	staged.total_bytes = 0;
	staged.row_count = 0;
	staged.failed = true;
}

void GpuColumnCache::AbortStagingLocked(const GpuTableRef &table) {
	auto it = staging_.find(table);
	if (it == staging_.end()) {
		return;
	}
	for (auto &kv : it->second) {
		FreeChunks(kv.second.chunks);
		staged_bytes_ -= kv.second.total_bytes;
	}
	staging_.erase(it);
}

void GpuColumnCache::BeginStaging(const GpuTableRef &table) {
	EnsureInitialized();
	std::lock_guard<std::mutex> guard(mutex_);
	AbortStagingLocked(table);
	if (budget_bytes_ == 0) {
		return;
	}
	staging_[table];
}

bool GpuColumnCache::StageChunk(const GpuTableRef &table, const std::string &column_name,
                                const GpuColumnChunk &chunk) {
	std::lock_guard<std::mutex> guard(mutex_);
	auto tit = staging_.find(table);
	if (tit == staging_.end()) {
		return false;
	}
	StagedColumn &staged = tit->second[column_name];
	if (staged.failed) {
		return false;
	}
	if (!staged.type_known) {
		staged.type = chunk.type;
		staged.type_known = true;
	} else if (staged.type != chunk.type) {
		FailStagedColumnLocked(staged);
		return false;
	}
	const uint64_t need = chunk.TotalBytes();
	while (committed_bytes_ + staged_bytes_ + need > budget_bytes_ && !lru_order_.empty()) {
		EvictLocked(lru_order_.back());
	}
	if (committed_bytes_ + staged_bytes_ + need > budget_bytes_) {
		FailStagedColumnLocked(staged);
		return false;
	}
	staged.chunks.push_back(chunk);
	staged.row_count += chunk.row_count;
	staged.total_bytes += need;
	staged_bytes_ += need;
	return true;
}

void GpuColumnCache::CommitStaging(const GpuTableRef &table, uint64_t scanned_rows) {
	std::lock_guard<std::mutex> guard(mutex_);
	auto tit = staging_.find(table);
	if (tit == staging_.end()) {
		return;
	}
	const uint64_t gen = generation_[table];
	for (auto &kv : tit->second) {
		StagedColumn &staged = kv.second;
		if (staged.failed || staged.chunks.empty() || scanned_rows == 0 || staged.row_count != scanned_rows) {
			FreeChunks(staged.chunks);
			staged_bytes_ -= staged.total_bytes;
			continue;
		}
		SlotKey slot {table, kv.first};
		EvictLocked(slot);
		CachedColumn col;
		col.chunks = std::move(staged.chunks);
		col.row_count = staged.row_count;
		col.total_bytes = staged.total_bytes;
		col.generation = gen;
		staged_bytes_ -= staged.total_bytes;
		committed_bytes_ += col.total_bytes;
		lru_order_.push_front(slot);
		entries_[slot] = {std::move(col), lru_order_.begin()};
	}
	staging_.erase(tit);
}

void GpuColumnCache::AbortStaging(const GpuTableRef &table) {
	std::lock_guard<std::mutex> guard(mutex_);
	AbortStagingLocked(table);
}

std::vector<uint64_t> GpuColumnCache::ReplayableChunkRows(const GpuTableRef &table,
                                                          const std::vector<std::string> &column_names) {
	std::lock_guard<std::mutex> guard(mutex_);
	std::vector<uint64_t> rows;
	if (column_names.empty()) {
		return rows;
	}
	for (size_t i = 0; i < column_names.size(); i++) {
		auto it = entries_.find(SlotKey {table, column_names[i]});
		if (it == entries_.end() || it->second.first.chunks.empty()) {
			return {};
		}
		const auto &chunks = it->second.first.chunks;
		if (i == 0) {
			for (auto &c : chunks) {
				rows.push_back(c.row_count);
			}
			continue;
		}
		if (chunks.size() != rows.size()) {
			return {};
		}
		for (size_t c = 0; c < chunks.size(); c++) {
			if (chunks[c].row_count != rows[c]) {
				return {};
			}
		}
	}
	for (auto &name : column_names) {
		TouchLocked(SlotKey {table, name});
	}
	return rows;
}

bool GpuColumnCache::TryGetChunk(const GpuTableRef &table, const std::string &column_name, uint64_t chunk_index,
                                 GpuColumnChunk &out_chunk) {
	std::lock_guard<std::mutex> guard(mutex_);
	SlotKey slot {table, column_name};
	auto it = entries_.find(slot);
	if (it == entries_.end() || chunk_index >= it->second.first.chunks.size()) {
		return false;
	}
	out_chunk = it->second.first.chunks[(size_t)chunk_index];
	TouchLocked(slot);
	return true;
}

void GpuColumnCache::InvalidateTable(const GpuTableRef &table) {
	std::lock_guard<std::mutex> guard(mutex_);
	AbortStagingLocked(table);
	EvictTableLocked(table);
	generation_[table]++;
}

void GpuColumnCache::InvalidateAll() {
	std::lock_guard<std::mutex> guard(mutex_);
	std::vector<GpuTableRef> tables;
	for (auto &kv : staging_) {
		tables.push_back(kv.first);
	}
	for (auto &t : tables) {
		AbortStagingLocked(t);
	}
	std::vector<SlotKey> slots;
	for (auto &kv : entries_) {
		slots.push_back(kv.first);
	}
	for (auto &s : slots) {
		EvictLocked(s);
	}
	for (auto &kv : generation_) {
		kv.second++;
	}
}

uint64_t GpuColumnCache::CurrentBytes() const {
	std::lock_guard<std::mutex> guard(mutex_);
	return committed_bytes_ + staged_bytes_;
}

uint64_t GpuColumnCache::TableGeneration(const GpuTableRef &table) const {
	std::lock_guard<std::mutex> guard(mutex_);
	auto it = generation_.find(table);
	return it == generation_.end() ? 0 : it->second;
}

} // namespace vector_gpu
