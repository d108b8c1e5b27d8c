// gpu_column_cache.cpp -- the device-resident column cache (the interface of
// cuda_engine/include/gpu_column_cache.hpp: staged scans, exact-row-count commit, replay, invalidation).
//
// State is one record per table: its generation, its committed ("resident") columns and the columns of an open
// scan ("pending"). Recency is a use clock: every resident column carries the tick of its last use and sits in
// an ordered index under that tick, so the eviction victim is always the index's first entry. Chunk memory
// belongs to the engine's device pool once staged, and goes back to it on eviction.
#include "../../../include/gpu_column_cache.hpp"

#include "../host/device_pool.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

namespace vector_gpu {

namespace {
constexpr uint64_t kWorkingReserve = 24ull << 30; // HBM left free for operator intermediates

void Release(std::vector<GpuColumnChunk> &chunks) {
	auto &pool = vgpu::DevicePool::Instance();
	for (auto &c : chunks) {
		pool.FreeAny(c.device_data);
		pool.FreeAny(c.device_validity);
	}
	chunks.clear();
}

bool EnvNumber(const char *name, double &out) {
	const char *text = std::getenv(name);
	if (text == nullptr || *text == '\0') {
		return false;
	}
	char *end = nullptr;
	const double v = std::strtod(text, &end);
	if (end == text || !(v > 0.0)) {
		return false;
	}
	out = v;
	return true;
}
} // namespace

GpuColumnCache &GpuColumnCache::Instance() {
	static GpuColumnCache *cache = new GpuColumnCache(); // never destroyed: outlives every plan
	return *cache;
}

GpuColumnCache::GpuColumnCache() = default;
GpuColumnCache::~GpuColumnCache() = default;

GpuColumnCache::Key GpuColumnCache::KeyOf(const GpuTableRef &table) {
	Key key;
	key.reserve(table.catalog.size() + table.schema.size() + table.table.size() + 2);
	key.append(table.catalog).push_back('\0');
	key.append(table.schema).push_back('\0');
	key.append(table.table);
	return key;
}

void GpuColumnCache::EnsureInitialized(double fraction_of_vram) {
	std::lock_guard<std::mutex> lock(mutex_);
	EnsureInitializedLocked(fraction_of_vram);
}

void GpuColumnCache::EnsureInitializedLocked(double fraction_of_vram) {
	if (initialized_) {
		return;
	}
	initialized_ = true;
	double mb = 0.0, fraction = 0.0;
	if (EnvNumber("VECTOR_GPU_CACHE_MB", mb)) {
		budget_bytes_ = static_cast<uint64_t>(mb) << 20;
		return;
	}
	if (EnvNumber("VECTOR_GPU_CACHE_FRACTION", fraction) && fraction <= 1.0) {
		fraction_of_vram = fraction;
	} else if (!(fraction_of_vram > 0.0 && fraction_of_vram <= 1.0)) {
		fraction_of_vram = 0.75;
	}
	size_t free_bytes = 0, total_bytes = 0;
	if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) {
		(void)hipGetLastError();
		budget_bytes_ = 0; // no device: nothing is ever staged
		return;
	}
	const uint64_t by_fraction = static_cast<uint64_t>(static_cast<double>(free_bytes) * fraction_of_vram);
	const uint64_t by_reserve = free_bytes > kWorkingReserve ? free_bytes - kWorkingReserve : 0;
	budget_bytes_ = std::min(by_fraction, by_reserve);
}

void GpuColumnCache::UseLocked(const Key &key, const std::string &column, Resident &r) {
	recency_.erase(r.last_use);
	r.last_use = ++use_clock_;
	recency_.emplace(r.last_use, std::make_pair(key, column));
}

void GpuColumnCache::DropResidentLocked(Table &table, std::map<std::string, Resident>::iterator it) {
	Resident &r = it->second;
	recency_.erase(r.last_use);
	resident_bytes_ -= r.bytes;
	Release(r.chunks);
	table.resident.erase(it);
}

void GpuColumnCache::DropPendingLocked(Table &table) {
	for (auto &kv : table.pending) {
		pending_bytes_ -= kv.second.bytes;
		Release(kv.second.chunks);
	}
	table.pending.clear();
	table.staging = false;
}

bool GpuColumnCache::MakeRoomLocked(uint64_t need) {
	while (resident_bytes_ + pending_bytes_ + need > budget_bytes_ && !recency_.empty()) {
		const auto victim = recency_.begin()->second; // least recently used column
		Table &t = tables_[victim.first];
		DropResidentLocked(t, t.resident.find(victim.second));
	}
	return resident_bytes_ + pending_bytes_ + need <= budget_bytes_;
}

void GpuColumnCache::BeginStaging(const GpuTableRef &table) {
	std::lock_guard<std::mutex> lock(mutex_);
	EnsureInitializedLocked(0.75);
	Table &t = tables_[KeyOf(table)];
	DropPendingLocked(t);
	t.staging = budget_bytes_ > 0;
}

bool GpuColumnCache::StageChunk(const GpuTableRef &table, const std::string &column_name,
                                const GpuColumnChunk &chunk) {
	std::lock_guard<std::mutex> lock(mutex_);
	auto tit = tables_.find(KeyOf(table));
	if (tit == tables_.end() || !tit->second.staging) {
		return false;
	}
	Pending &p = tit->second.pending[column_name];
	if (p.dead) {
		return false;
	}
	const bool type_ok = !p.typed || p.type == chunk.type;
	p.type = chunk.type;
	p.typed = true;
	if (!type_ok || !MakeRoomLocked(chunk.TotalBytes())) {
		// this scan cannot complete the column: give back what it staged, refuse the rest
		pending_bytes_ -= p.bytes;
		Release(p.chunks);
		p.bytes = 0;
		p.rows = 0;
		p.dead = true;
		return false;
	}
	p.chunks.push_back(chunk);
	p.rows += chunk.row_count;
	p.bytes += chunk.TotalBytes();
	pending_bytes_ += chunk.TotalBytes();
	return true;
}

void GpuColumnCache::CommitStaging(const GpuTableRef &table, uint64_t scanned_rows) {
	std::lock_guard<std::mutex> lock(mutex_);
	const Key key = KeyOf(table);
	auto tit = tables_.find(key);
	if (tit == tables_.end() || !tit->second.staging) {
		return;
	}
	Table &t = tit->second;
	for (auto &kv : t.pending) {
		Pending &p = kv.second;
		pending_bytes_ -= p.bytes;
		if (p.dead || p.chunks.empty() || scanned_rows == 0 || p.rows != scanned_rows) {
			Release(p.chunks);
			continue;
		}
		auto old = t.resident.find(kv.first);
		if (old != t.resident.end()) {
			DropResidentLocked(t, old);
		}
		Resident &r = t.resident[kv.first];
		r.chunks = std::move(p.chunks);
		r.rows = p.rows;
		r.bytes = p.bytes;
		resident_bytes_ += r.bytes;
		r.last_use = ++use_clock_;
		recency_.emplace(r.last_use, std::make_pair(key, kv.first));
	}
	t.pending.clear();
	t.staging = false;
}

void GpuColumnCache::AbortStaging(const GpuTableRef &table) {
	std::lock_guard<std::mutex> lock(mutex_);
	auto tit = tables_.find(KeyOf(table));
	if (tit != tables_.end()) {
		DropPendingLocked(tit->second);
	}
}

std::vector<uint64_t> GpuColumnCache::ReplayableChunkRows(const GpuTableRef &table,
                                                          const std::vector<std::string> &column_names) {
	std::lock_guard<std::mutex> lock(mutex_);
	const Key key = KeyOf(table);
	auto tit = tables_.find(key);
	if (column_names.empty() || tit == tables_.end()) {
		return {};
	}
	std::vector<Resident *> cols;
	for (const auto &name : column_names) {
		auto it = tit->second.resident.find(name);
		if (it == tit->second.resident.end() || it->second.chunks.empty()) {
			return {};
		}
		cols.push_back(&it->second);
	}
	std::vector<uint64_t> rows;
	for (const auto &c : cols.front()->chunks) {
		rows.push_back(c.row_count);
	}
	for (const Resident *r : cols) {
		const bool same = r->chunks.size() == rows.size() &&
		                  std::equal(rows.begin(), rows.end(), r->chunks.begin(),
		                             [](uint64_t n, const GpuColumnChunk &c) { return n == c.row_count; });
		if (!same) {
			return {};
		}
	}
	for (size_t i = 0; i < cols.size(); i++) {
		UseLocked(key, column_names[i], *cols[i]);
	}
	return rows;
}

bool GpuColumnCache::TryGetChunk(const GpuTableRef &table, const std::string &column_name, uint64_t chunk_index,
                                 GpuColumnChunk &out_chunk) {
	std::lock_guard<std::mutex> lock(mutex_);
	const Key key = KeyOf(table);
	auto tit = tables_.find(key);
	if (tit == tables_.end()) {
		return false;
	}
	auto it = tit->second.resident.find(column_name);
	if (it == tit->second.resident.end() || chunk_index >= it->second.chunks.size()) {
		return false;
	}
	out_chunk = it->second.chunks[static_cast<size_t>(chunk_index)];
	UseLocked(key, column_name, it->second);
	return true;
}

void GpuColumnCache::InvalidateTable(const GpuTableRef &table) {
	std::lock_guard<std::mutex> lock(mutex_);
	Table &t = tables_[KeyOf(table)];
	DropPendingLocked(t);
	while (!t.resident.empty()) {
		DropResidentLocked(t, t.resident.begin());
	}
	t.generation++;
}

void GpuColumnCache::InvalidateAll() {
	std::lock_guard<std::mutex> lock(mutex_);
	for (auto &kv : tables_) {
		Table &t = kv.second;
		DropPendingLocked(t);
		while (!t.resident.empty()) {
			DropResidentLocked(t, t.resident.begin());
		}
		t.generation++;
	}
}

uint64_t GpuColumnCache::CurrentBytes() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return resident_bytes_ + pending_bytes_;
}

uint64_t GpuColumnCache::TableGeneration(const GpuTableRef &table) const {
	std::lock_guard<std::mutex> lock(mutex_);
	auto it = tables_.find(KeyOf(table));
	return it == tables_.end() ? 0 : it->second.generation;
}

} // namespace vector_gpu
