// gpu_memory_pool.cpp -- cyclic ring of page-locked batch slots (contract of
// cuda_engine/src/gpu_memory_pool.cpp).
#include "../../../include/gpu_memory_pool.hpp"
#include "../../../include/gpu_logger.hpp"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace vector_gpu {

GpuMemoryPoolSlot::GpuMemoryPoolSlot(GpuMemoryPoolSlot &&other) noexcept
    : pool_(other.pool_), index_(other.index_), data_(other.data_), capacity_(other.capacity_) {
	other.pool_ = nullptr;
	other.index_ = -1;
	other.data_ = nullptr;
	other.capacity_ = 0;
}

GpuMemoryPoolSlot &GpuMemoryPoolSlot::operator=(GpuMemoryPoolSlot &&other) noexcept {
	if (this != &other) {
		if (pool_ != nullptr) {
			pool_->Release(index_);
		}
		pool_ = other.pool_;
		index_ = other.index_;
		data_ = other.data_;
		capacity_ = other.capacity_;
		other.pool_ = nullptr;
		other.index_ = -1;
		other.data_ = nullptr;
		other.capacity_ = 0;
	}
	return *this;
}

GpuMemoryPoolSlot::~GpuMemoryPoolSlot() {
	if (pool_ != nullptr) {
		pool_->Release(index_);
	}
}

GpuMemoryPool &GpuMemoryPool::Instance() {
	static GpuMemoryPool instance;
	return instance;
}

GpuMemoryPool::GpuMemoryPool() {
	for (size_t i = 0; i < kSlotCount; i++) {
		void *ptr = nullptr;
		hipError_t status = hipHostMalloc(&ptr, kSlotBytes, hipHostMallocPortable);
		if (status != hipSuccess) {
			(void)hipGetLastError();
			for (size_t j = 0; j < i; j++) {
				(void)hipHostFree(slots_[j].data);
				slots_[j].data = nullptr;
			}
			throw std::runtime_error("GpuMemoryPool: hipHostMalloc(" + std::to_string(kSlotBytes) +
			                         ", hipHostMallocPortable) failed for slot " + std::to_string(i) + ": " +
			                         hipGetErrorString(status));
		}
		slots_[i].data = ptr;
	}
}

GpuMemoryPool::~GpuMemoryPool() {
	for (auto &slot : slots_) {
		if (slot.data != nullptr) {
			(void)hipHostFree(slot.data);
		}
	}
}

GpuMemoryPoolSlot GpuMemoryPool::AcquireNext() {
	std::unique_lock<std::mutex> lock(mutex_);
	const size_t index = next_index_;
	cv_.wait(lock, [&] { return !slots_[index].in_use; });
	slots_[index].in_use = true;
	next_index_ = (index + 1) % kSlotCount;
	if (GpuLogger::Instance().Enabled(LogLevel::DEBUG)) {
		size_t active = 0;
		for (const auto &s : slots_) {
			active += s.in_use ? 1 : 0;
		}
		VGPU_LOG(LogLevel::DEBUG, "gpu_memory_pool",
		         "\"active_slots\":" + std::to_string(active) + ",\"total_slots\":" + std::to_string(kSlotCount));
	}
	return GpuMemoryPoolSlot(this, static_cast<int>(index), slots_[index].data, kSlotBytes);
}

void GpuMemoryPool::ResetCursorForTesting() {
	std::lock_guard<std::mutex> guard(mutex_);
	next_index_ = 0;
}

void GpuMemoryPool::Release(int index) {
	{
		std::lock_guard<std::mutex> guard(mutex_);
		slots_[static_cast<size_t>(index)].in_use = false;
	}
	// Waiters block on one specific index, so every waiter must re-check.
	cv_.notify_all();
}

} // namespace vector_gpu
