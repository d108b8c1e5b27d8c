// rmm_pool.cpp -- device-memory budget oracle and best-fit arena (the interface of
// cuda_engine/include/rmm_pool.hpp).
//
// The budget is a fraction of the HBM free at first use. Routing asks HasHeadroomFor(working set) before it
// sends a plan to the device; the answer is whether one hole of the arena's free space could hold it. The arena
// itself (one hipMalloc of the whole budget) is only reserved by the first Allocate().
#include "../../../include/rmm_pool.hpp"
#include "../../../include/gpu_logger.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace vector_gpu {

namespace {
constexpr size_t kAlign = 256;

size_t Aligned(size_t n) {
	return (n + kAlign - 1) / kAlign * kAlign;
}

double FractionFromEnv(double fallback) {
	const char *text = std::getenv("VECTOR_GPU_VRAM_FRACTION");
	if (text == nullptr || *text == '\0') {
		return fallback;
	}
	char *end = nullptr;
	const double v = std::strtod(text, &end);
	return (end != text && v > 0.0 && v <= 1.0) ? v : fallback;
}
} // namespace

RmmPool &RmmPool::Instance() {
	static RmmPool *pool = new RmmPool(); // never destroyed: the runtime may be gone at static teardown
	return *pool;
}

RmmPool::RmmPool() = default;

RmmPool::~RmmPool() {
	if (arena_ != nullptr) {
		(void)hipFree(arena_);
	}
}

void RmmPool::AddHoleLocked(size_t offset, size_t size) {
	// merge with the hole that ends here and the hole that starts right after
	auto next = holes_.lower_bound(offset);
	if (next != holes_.begin()) {
		auto prev = std::prev(next);
		if (prev->first + prev->second == offset) {
			offset = prev->first;
			size += prev->second;
			RemoveHoleLocked(prev->first, prev->second);
		}
	}
	next = holes_.find(offset + size);
	if (next != holes_.end()) {
		size += next->second;
		RemoveHoleLocked(next->first, next->second);
	}
	holes_.emplace(offset, size);
	holes_by_size_.emplace(size, offset);
}

void RmmPool::RemoveHoleLocked(size_t offset, size_t size) {
	holes_.erase(offset);
	holes_by_size_.erase({size, offset});
}

void RmmPool::SetBudgetLocked(size_t bytes) {
	total_pool_bytes_ = bytes;
	holes_.clear();
	holes_by_size_.clear();
	taken_.clear();
	taken_bytes_ = 0;
	AddHoleLocked(0, bytes);
	initialized_ = true;
}

void RmmPool::EnsureInitialized(double fraction_of_vram) {
	std::lock_guard<std::mutex> lock(mutex_);
	if (initialized_) {
		return;
	}
	fraction_of_vram = FractionFromEnv(fraction_of_vram);
	if (!(fraction_of_vram > 0.0 && fraction_of_vram <= 1.0)) {
		throw std::invalid_argument("RmmPool::EnsureInitialized: fraction_of_vram must be in (0, 1]");
	}
	size_t free_bytes = 0, total_bytes = 0;
	const hipError_t status = hipMemGetInfo(&free_bytes, &total_bytes);
	if (status != hipSuccess) {
		(void)hipGetLastError();
		throw std::runtime_error(std::string("RmmPool: cannot query device memory: ") + hipGetErrorString(status));
	}
	const size_t budget = Aligned(static_cast<size_t>(static_cast<double>(free_bytes) * fraction_of_vram));
	if (budget == 0) {
		throw std::runtime_error("RmmPool: the budget rounds to 0 bytes");
	}
	SetBudgetLocked(budget);
	VGPU_LOG(LogLevel::INFO, "vram_budget",
	         "\"budget_mib\":" + std::to_string(budget >> 20) + ",\"hbm_free_mib\":" + std::to_string(free_bytes >> 20) +
	             ",\"hbm_total_mib\":" + std::to_string(total_bytes >> 20) +
	             ",\"fraction\":" + std::to_string(fraction_of_vram));
}

void RmmPool::EnsureInitializedExactBytesForTesting(size_t bytes) {
	std::lock_guard<std::mutex> lock(mutex_);
	if (initialized_) {
		return;
	}
	const size_t budget = Aligned(bytes);
	if (budget == 0) {
		throw std::runtime_error("RmmPool: the budget rounds to 0 bytes");
	}
	SetBudgetLocked(budget);
	if (!ReserveArenaLocked()) {
		initialized_ = false;
		total_pool_bytes_ = 0;
		throw std::runtime_error("RmmPool: cannot reserve an arena of " + std::to_string(budget) + " bytes");
	}
}

bool RmmPool::ReserveArenaLocked() {
	if (arena_ != nullptr) {
		return true;
	}
	if (total_pool_bytes_ == 0 || hipMalloc(&arena_, total_pool_bytes_) != hipSuccess) {
		(void)hipGetLastError();
		arena_ = nullptr;
		return false;
	}
	return true;
}

bool RmmPool::HasHeadroomFor(size_t estimated_bytes) const {
	std::lock_guard<std::mutex> lock(mutex_);
	if (!initialized_ || estimated_bytes == 0 || estimated_bytes > total_pool_bytes_ || holes_by_size_.empty()) {
		return false;
	}
	return holes_by_size_.rbegin()->first >= Aligned(estimated_bytes);
}

void *RmmPool::Allocate(size_t bytes) {
	if (bytes == 0) {
		return nullptr;
	}
	std::lock_guard<std::mutex> lock(mutex_);
	if (!initialized_ || bytes > total_pool_bytes_ || !ReserveArenaLocked()) {
		return nullptr;
	}
	const size_t need = Aligned(bytes);
	auto fit = holes_by_size_.lower_bound({need, 0});
	if (fit == holes_by_size_.end()) {
		return nullptr;
	}
	const size_t size = fit->first, offset = fit->second;
	RemoveHoleLocked(offset, size);
	if (size > need) {
		AddHoleLocked(offset + need, size - need);
	}
	taken_.emplace(offset, need);
	taken_bytes_ += need;
	return static_cast<unsigned char *>(arena_) + offset;
}

void RmmPool::Free(void *ptr, size_t /*bytes: the pool tracks each block's size*/) {
	if (ptr == nullptr) {
		return;
	}
	std::lock_guard<std::mutex> lock(mutex_);
	if (arena_ == nullptr) {
		return;
	}
	const auto addr = reinterpret_cast<uintptr_t>(ptr), base = reinterpret_cast<uintptr_t>(arena_);
	if (addr < base || addr >= base + total_pool_bytes_) {
		return; // not from this arena
	}
	auto it = taken_.find(static_cast<size_t>(addr - base));
	if (it == taken_.end()) {
		return; // interior pointer or already freed
	}
	const size_t offset = it->first, size = it->second;
	taken_.erase(it);
	taken_bytes_ -= size;
	AddHoleLocked(offset, size);
}

size_t RmmPool::AllocatedBytes() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return taken_bytes_;
}

void RmmPool::ResetForTesting() {
	std::lock_guard<std::mutex> lock(mutex_);
	if (arena_ != nullptr) {
		(void)hipFree(arena_);
		arena_ = nullptr;
	}
	total_pool_bytes_ = 0;
	holes_.clear();
	holes_by_size_.clear();
	taken_.clear();
	taken_bytes_ = 0;
	initialized_ = false;
}

} // namespace vector_gpu
