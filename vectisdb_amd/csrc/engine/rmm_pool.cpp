// rmm_pool.cpp -- budget oracle + lazily reserved first-fit arena (contract of
// cuda_engine/src/rmm_pool.cpp and cuda_engine/include/rmm_pool.hpp).
#include "../../../include/rmm_pool.hpp"
#include "../../../include/gpu_logger.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace vector_gpu {

namespace {
constexpr size_t kAlignment = 256;

size_t AlignUp(size_t n) {
	return (n + kAlignment - 1) & ~(kAlignment - 1);
}
} // namespace

RmmPool &RmmPool::Instance() {
	static RmmPool *pool = new RmmPool();
	return *pool;
}

RmmPool::RmmPool() = default;

RmmPool::~RmmPool() {
	if (arena_ != nullptr) {
		(void)hipFree(arena_);
	}
}

void RmmPool::ResetBlocksLocked(size_t total) {
	total_pool_bytes_ = total;
	free_blocks_.clear();
	free_blocks_.emplace(0, total);
	live_blocks_.clear();
	initialized_ = true;
}

void RmmPool::AllocateArenaLocked(size_t requested_bytes) {
	size_t aligned = AlignUp(requested_bytes);
	if (aligned == 0) {
		throw std::runtime_error("RmmPool: computed pool size is 0 bytes");
	}
	hipError_t status = hipMalloc(&arena_, aligned);
	if (status != hipSuccess) {
		(void)hipGetLastError();
		arena_ = nullptr;
		throw std::runtime_error("RmmPool: hipMalloc(" + std::to_string(aligned) + ") failed: " +
		                         hipGetErrorString(status));
	}
	ResetBlocksLocked(aligned);
}

void RmmPool::EnsureInitialized(double fraction_of_vram) {
	std::lock_guard<std::mutex> guard(mutex_);
	if (initialized_) {
		return;
	}
	if (const char *env = std::getenv("VECTOR_GPU_VRAM_FRACTION")) {
		char *end = nullptr;
		double parsed = std::strtod(env, &end);
		if (end != env && parsed > 0.0 && parsed <= 1.0) {
			fraction_of_vram = parsed;
		}
	}
	if (!(fraction_of_vram > 0.0 && fraction_of_vram <= 1.0)) {
		throw std::invalid_argument("RmmPool::EnsureInitialized: fraction_of_vram must be in (0, 1]");
	}
	size_t free_bytes = 0, total_bytes = 0;
	hipError_t status = hipMemGetInfo(&free_bytes, &total_bytes);
	if (status != hipSuccess) {
		(void)hipGetLastError();
		throw std::runtime_error(std::string("RmmPool: hipMemGetInfo failed: ") + hipGetErrorString(status));
	}
	size_t aligned = AlignUp(static_cast<size_t>(static_cast<double>(free_bytes) * fraction_of_vram));
	if (aligned == 0) {
		throw std::runtime_error("RmmPool: computed pool size is 0 bytes");
	}
	ResetBlocksLocked(aligned);
	VGPU_LOG(LogLevel::INFO, "vram_budget",
	         "\"locked_mib\":" + std::to_string(aligned >> 20) + ",\"free_mib\":" + std::to_string(free_bytes >> 20) +
	             ",\"total_mib\":" + std::to_string(total_bytes >> 20) + ",\"fraction\":" +
	             std::to_string(fraction_of_vram) + ",\"one_time_snapshot\":true");
}

bool RmmPool::EnsureArenaLocked() {
	if (arena_ != nullptr) {
		return true;
	}
	if (total_pool_bytes_ == 0) {
		return false;
	}
	if (hipMalloc(&arena_, total_pool_bytes_) != hipSuccess) {
		(void)hipGetLastError();
		arena_ = nullptr;
		return false;
	}
	return true;
}

void RmmPool::EnsureInitializedExactBytesForTesting(size_t bytes) {
	std::lock_guard<std::mutex> guard(mutex_);
	if (initialized_) {
		return;
	}
	AllocateArenaLocked(bytes);
}

bool RmmPool::HasHeadroomFor(size_t estimated_bytes) const {
	std::lock_guard<std::mutex> guard(mutex_);
	if (!initialized_ || estimated_bytes == 0 || estimated_bytes > total_pool_bytes_) {
		return false;
	}
	size_t aligned = AlignUp(estimated_bytes);
	for (auto &kv : free_blocks_) {
		if (kv.second >= aligned) {
			return true;
		}
	}
	return false;
}

void *RmmPool::Allocate(size_t bytes) {
	if (bytes == 0) {
		return nullptr;
	}
	std::lock_guard<std::mutex> guard(mutex_);
	if (!initialized_ || bytes > total_pool_bytes_ || !EnsureArenaLocked()) {
		return nullptr;
	}
	size_t aligned = AlignUp(bytes);
	for (auto it = free_blocks_.begin(); it != free_blocks_.end(); ++it) {
		if (it->second < aligned) {
			continue;
		}
		size_t offset = it->first;
		size_t remaining = it->second - aligned;
		free_blocks_.erase(it);
		if (remaining > 0) {
			free_blocks_.emplace(offset + aligned, remaining);
		}
		live_blocks_[offset] = aligned;
		return static_cast<uint8_t *>(arena_) + offset;
	}
	return nullptr;
}

void RmmPool::Free(void *ptr, size_t bytes) {
	if (ptr == nullptr || bytes == 0) {
		return;
	}
	std::lock_guard<std::mutex> guard(mutex_);
	if (!initialized_ || arena_ == nullptr) {
		return;
	}
	auto *p = static_cast<uint8_t *>(ptr);
	auto *base = static_cast<uint8_t *>(arena_);
	if (p < base || p >= base + total_pool_bytes_) {
		return;
	}
	size_t offset = static_cast<size_t>(p - base);
	auto live = live_blocks_.find(offset);
	if (live == live_blocks_.end()) {
		return;
	}
	size_t size = live->second;
	live_blocks_.erase(live);
	auto it = free_blocks_.emplace(offset, size).first;
	auto next = std::next(it);
	if (next != free_blocks_.end() && it->first + it->second == next->first) {
		it->second += next->second;
		free_blocks_.erase(next);
	}
	if (it != free_blocks_.begin()) {
		auto prev = std::prev(it);
		if (prev->first + prev->second == it->first) {
			prev->second += it->second;
			free_blocks_.erase(it);
		}
	}
}

size_t RmmPool::AllocatedBytes() const {
	std::lock_guard<std::mutex> guard(mutex_);
	size_t free_total = 0;
	for (auto &kv : free_blocks_) {
		free_total += kv.second;
	}
	return total_pool_bytes_ - free_total;
}

void RmmPool::ResetForTesting() {
	std::lock_guard<std::mutex> guard(mutex_);
	if (arena_ != nullptr) {
		(void)hipFree(arena_);
		arena_ = nullptr;
	}
	total_pool_bytes_ = 0;
	free_blocks_.clear();
	live_blocks_.clear();
	initialized_ = false;
}

} // namespace vector_gpu
