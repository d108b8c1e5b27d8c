// pinned_buffer_pool.cpp -- page-locked staging slots (semantics of
// cuda_engine/src/pinned_buffer_pool.cpp: lazy growth to kMaxSlots, grow-only slot capacity, blocking
// Acquire at the cap).
#include "../../../include/pinned_buffer_pool.hpp"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace vector_gpu {

PinnedBufferSlot::PinnedBufferSlot(PinnedBufferSlot &&other) noexcept
    : pool_(other.pool_), index_(other.index_), data_(other.data_) {
	other.pool_ = nullptr;
	other.index_ = -1;
	other.data_ = nullptr;
}

PinnedBufferSlot &PinnedBufferSlot::operator=(PinnedBufferSlot &&other) noexcept {
	if (this != &other) {
		if (pool_ != nullptr) {
			pool_->Release(index_);
		}
		pool_ = other.pool_;
		index_ = other.index_;
		data_ = other.data_;
		other.pool_ = nullptr;
		other.index_ = -1;
		other.data_ = nullptr;
	}
	return *this;
}

PinnedBufferSlot::~PinnedBufferSlot() {
	if (pool_ != nullptr) {
		pool_->Release(index_);
	}
}

PinnedBufferPool::PinnedBufferPool() = default;

PinnedBufferPool::~PinnedBufferPool() {
	for (auto &slot : slots_) {
		if (slot.data != nullptr) {
			(void)hipHostFree(slot.data);
		}
	}
}

PinnedBufferPool &PinnedBufferPool::Instance() {
	// Leaked on purpose: freeing page-locked memory during static destruction races the runtime's teardown.
	static PinnedBufferPool *pool = new PinnedBufferPool();
	return *pool;
}

PinnedBufferSlot PinnedBufferPool::Acquire(size_t bytes) {
	std::unique_lock<std::mutex> lock(mutex_);
	int chosen = -1;
	cv_.wait(lock, [&] {
		// Prefer a free slot that is already large enough, then any free slot, then a new slot.
		int any = -1;
		for (size_t i = 0; i < slots_.size(); i++) {
			if (slots_[i].in_use) {
				continue;
			}
			if (slots_[i].capacity >= bytes) {
				chosen = static_cast<int>(i);
				return true;
			}
			if (any < 0) {
				any = static_cast<int>(i);
			}
		}
		if (any >= 0) {
			chosen = any;
			return true;
		}
		if (slots_.size() < kMaxSlots) {
			slots_.emplace_back();
			chosen = static_cast<int>(slots_.size() - 1);
			return true;
		}
		return false;
	});
	auto &slot = slots_[static_cast<size_t>(chosen)];
	if (bytes > slot.capacity) {
		void *grown = nullptr;
		hipError_t status = hipHostMalloc(&grown, bytes, hipHostMallocDefault);
		if (status != hipSuccess) {
			(void)hipGetLastError();
			throw std::runtime_error(std::string("PinnedBufferPool::Acquire: hipHostMalloc(") + std::to_string(bytes) +
			                         ") failed: " + hipGetErrorString(status));
		}
		if (slot.data != nullptr) {
			(void)hipHostFree(slot.data);
		}
		slot.data = grown;
		slot.capacity = bytes;
	}
	slot.in_use = true;
	return PinnedBufferSlot(this, chosen, slot.data);
}

void PinnedBufferPool::Release(int index) {
	{
		std::lock_guard<std::mutex> guard(mutex_);
		slots_[static_cast<size_t>(index)].in_use = false;
	}
	cv_.notify_one();
}

} // namespace vector_gpu
