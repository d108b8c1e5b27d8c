// gpu_memory_pool.cpp -- the stream pipeline's ring of page-locked batch slots (the interface of
// cuda_engine/include/gpu_memory_pool.hpp).
//
// The ring is one page-locked region: one hipHostMalloc call and one DMA mapping for all four slots. Slots go
// out through a ticket turnstile: the n-th caller of AcquireNext holds ticket n and gets slot n % kSlotCount
// as soon as that slot has come back from its previous holder, so concurrent producers are served strictly
// in arrival order and never contend for the same slot.
#include "../../../include/gpu_memory_pool.hpp"
#include "../../../include/gpu_logger.hpp"

#include <hip/hip_runtime.h>

#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>

namespace vector_gpu {

GpuMemoryPoolSlot::GpuMemoryPoolSlot(GpuMemoryPoolSlot &&other) noexcept
    : pool_(std::exchange(other.pool_, nullptr)), index_(std::exchange(other.index_, -1)),
      data_(std::exchange(other.data_, nullptr)), capacity_(std::exchange(other.capacity_, 0)) {
}

GpuMemoryPoolSlot &GpuMemoryPoolSlot::operator=(GpuMemoryPoolSlot &&other) noexcept {
	GpuMemoryPoolSlot taken(std::move(other)); // `other` is empty from here on, even when it is *this
	std::swap(pool_, taken.pool_);
	std::swap(index_, taken.index_);
	std::swap(data_, taken.data_);
	std::swap(capacity_, taken.capacity_);
	return *this; // `taken` now holds what this slot held before and returns it to its ring
}

GpuMemoryPoolSlot::~GpuMemoryPoolSlot() {
	if (auto *pool = std::exchange(pool_, nullptr)) {
		pool->Return(index_);
	}
}

GpuMemoryPool &GpuMemoryPool::Instance() {
	static GpuMemoryPool ring;
	return ring;
}

GpuMemoryPool::GpuMemoryPool() {
	void *region = nullptr;
	const size_t bytes = kSlotCount * kSlotBytes;
	const hipError_t status = hipHostMalloc(&region, bytes, hipHostMallocPortable);
	if (status != hipSuccess || region == nullptr) {
		(void)hipGetLastError();
		throw std::runtime_error("GpuMemoryPool: cannot page-lock the " + std::to_string(bytes >> 20) +
		                         " MiB slot ring: " + hipGetErrorString(status));
	}
	region_ = static_cast<unsigned char *>(region);
}

GpuMemoryPool::~GpuMemoryPool() {
	if (region_ != nullptr) {
		(void)hipHostFree(region_);
	}
}

GpuMemoryPoolSlot GpuMemoryPool::AcquireNext() {
	std::unique_lock<std::mutex> lock(mutex_);
	const unsigned long long ticket = next_ticket_++;
	const size_t slot = static_cast<size_t>(ticket % kSlotCount);
	const unsigned long long round = ticket / kSlotCount;
	returned_.wait(lock, [&] { return returns_[slot] >= round; });
	VGPU_LOG(LogLevel::DEBUG, "gpu_memory_pool",
	         "\"ticket\":" + std::to_string(ticket) + ",\"slot\":" + std::to_string(slot) + ",\"outstanding\":" +
	             std::to_string(next_ticket_ - std::accumulate(returns_.begin(), returns_.end(), 0ull)));
	return GpuMemoryPoolSlot(this, static_cast<int>(slot), region_ + slot * kSlotBytes, kSlotBytes);
}

void GpuMemoryPool::ResetCursorForTesting() {
	std::lock_guard<std::mutex> lock(mutex_);
	// Start a fresh round: every slot counts as returned once per round issued so far.
	const unsigned long long rounds = (next_ticket_ + kSlotCount - 1) / kSlotCount;
	next_ticket_ = rounds * kSlotCount;
	returns_.fill(rounds);
}

void GpuMemoryPool::Return(int index) {
	{
		std::lock_guard<std::mutex> lock(mutex_);
		returns_[static_cast<size_t>(index)]++;
	}
	returned_.notify_all(); // each waiter waits for its own slot and round
}

} // namespace vector_gpu
