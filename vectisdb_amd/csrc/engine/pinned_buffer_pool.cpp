// pinned_buffer_pool.cpp -- size-classed cache of page-locked staging buffers (the engine-private role of
// cuda_engine/include/pinned_buffer_pool.hpp: staging for host <-> device copies of plan columns).
#include "../../../include/pinned_buffer_pool.hpp"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <utility>

namespace vector_gpu {

namespace {
size_t SizeClass(size_t bytes) {
	size_t c = PinnedBufferPool::kMinClassBytes;
	while (c < bytes) {
		c <<= 1;
	}
	return c;
}
} // namespace

PinnedBufferSlot::PinnedBufferSlot(PinnedBufferSlot &&other) noexcept
    : pool_(std::exchange(other.pool_, nullptr)), data_(std::exchange(other.data_, nullptr)),
      bytes_(std::exchange(other.bytes_, 0)) {
}

PinnedBufferSlot &PinnedBufferSlot::operator=(PinnedBufferSlot &&other) noexcept {
	PinnedBufferSlot taken(std::move(other));
	std::swap(pool_, taken.pool_);
	std::swap(data_, taken.data_);
	std::swap(bytes_, taken.bytes_);
	return *this; // the buffer this slot held before goes back with `taken`
}

PinnedBufferSlot::~PinnedBufferSlot() {
	if (auto *pool = std::exchange(pool_, nullptr)) {
		pool->Return(data_, bytes_);
	}
}

PinnedBufferPool &PinnedBufferPool::Instance() {
	// Never destroyed: releasing page-locked memory during static destruction races the runtime's teardown.
	static PinnedBufferPool *pool = new PinnedBufferPool();
	return *pool;
}

PinnedBufferPool::~PinnedBufferPool() {
	for (auto &cls : idle_) {
		for (void *p : cls.second) {
			(void)hipHostFree(p);
		}
	}
}

PinnedBufferSlot PinnedBufferPool::Acquire(size_t bytes) {
	const size_t cls = SizeClass(bytes);
	std::unique_lock<std::mutex> lock(mutex_);
	for (;;) {
		// the smallest idle buffer that covers the request
		auto fit = idle_.lower_bound(cls);
		while (fit != idle_.end() && fit->second.empty()) {
			++fit;
		}
		if (fit != idle_.end()) {
			void *p = fit->second.back();
			fit->second.pop_back();
			return PinnedBufferSlot(this, p, fit->first);
		}
		if (live_buffers_ >= kMaxSlots) {
			// at the cap: give back the largest idle buffer (all idle ones are too small), else wait
			auto largest = idle_.rbegin();
			while (largest != idle_.rend() && largest->second.empty()) {
				++largest;
			}
			if (largest == idle_.rend()) {
				returned_.wait(lock);
				continue;
			}
			(void)hipHostFree(largest->second.back());
			largest->second.pop_back();
			live_buffers_--;
			live_bytes_ -= largest->first;
		}
		void *p = nullptr;
		const hipError_t status = hipHostMalloc(&p, cls, hipHostMallocDefault);
		if (status != hipSuccess || p == nullptr) {
			(void)hipGetLastError();
			throw std::runtime_error("PinnedBufferPool::Acquire: cannot page-lock " + std::to_string(cls) +
			                         " bytes: " + hipGetErrorString(status));
		}
		live_buffers_++;
		live_bytes_ += cls;
		return PinnedBufferSlot(this, p, cls);
	}
}

void PinnedBufferPool::Return(void *data, size_t bytes) {
	{
		std::lock_guard<std::mutex> lock(mutex_);
		idle_[bytes].push_back(data);
	}
	returned_.notify_one();
}

size_t PinnedBufferPool::LiveBuffers() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return live_buffers_;
}

size_t PinnedBufferPool::LiveBytes() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return live_bytes_;
}

} // namespace vector_gpu
