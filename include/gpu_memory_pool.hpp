#pragma once
// gpu_memory_pool.hpp -- fixed ring of page-locked batch slots feeding the stream pipeline (drop-in for
// cuda_engine/include/gpu_memory_pool.hpp).
//
// kSlotCount slots of kSlotBytes each, carved from one page-locked region (hipHostMalloc, portable) the first
// time Instance() runs, and handed out in cyclic order. One slot per concurrently draining batch of the
// pipeline (kDepth = 3 in flight) plus one for the producer: GpuStreamPipeline static_asserts the
// relationship. A slot returns to the ring when its GpuMemoryPoolSlot is destroyed.

#include <array>
#include <condition_variable>
#include <cstddef>
#include <mutex>

namespace vector_gpu {

class GpuMemoryPoolSlot {
public:
	GpuMemoryPoolSlot() = default;
	GpuMemoryPoolSlot(const GpuMemoryPoolSlot &) = delete;
	GpuMemoryPoolSlot &operator=(const GpuMemoryPoolSlot &) = delete;
	GpuMemoryPoolSlot(GpuMemoryPoolSlot &&other) noexcept;
	GpuMemoryPoolSlot &operator=(GpuMemoryPoolSlot &&other) noexcept;
	~GpuMemoryPoolSlot();

	//! Page-locked host memory, exactly GpuMemoryPool::kSlotBytes long; nullptr when default-constructed
	//! or moved from.
	void *data() const {
		return data_;
	}
	size_t capacity() const {
		return capacity_;
	}
	bool valid() const {
		return data_ != nullptr;
	}

private:
	friend class GpuMemoryPool;
	GpuMemoryPoolSlot(class GpuMemoryPool *pool, int index, void *data, size_t capacity)
	    : pool_(pool), index_(index), data_(data), capacity_(capacity) {
	}
	class GpuMemoryPool *pool_ = nullptr;
	int index_ = -1;
	void *data_ = nullptr;
	size_t capacity_ = 0;
};

class GpuMemoryPool {
public:
	static constexpr size_t kSlotCount = 4;
	static constexpr size_t kSlotBytes = 350ull * 1024 * 1024;

	//! Page-locks the whole ring on first use; throws std::runtime_error when it cannot be page-locked, and
	//! the next call tries again.
	static GpuMemoryPool &Instance();

	GpuMemoryPool(const GpuMemoryPool &) = delete;
	GpuMemoryPool &operator=(const GpuMemoryPool &) = delete;

	//! Takes the next slot in cyclic order, blocking until its previous holder has returned it. Callers are
	//! served in arrival order (each takes a ticket; ticket t is slot t % kSlotCount).
	GpuMemoryPoolSlot AcquireNext();

	//! Test-only: the next AcquireNext takes slot 0 (no slot may be checked out).
	void ResetCursorForTesting();

private:
	GpuMemoryPool();
	~GpuMemoryPool();
	friend class GpuMemoryPoolSlot;
	void Return(int index);

	// One page-locked region of kSlotCount * kSlotBytes; slot i is its i-th kSlotBytes.
	unsigned char *region_ = nullptr;
	// Turnstile: ticket t may take slot t % kSlotCount once that slot has been returned t / kSlotCount times.
	std::mutex mutex_;
	std::condition_variable returned_;
	unsigned long long next_ticket_ = 0;
	std::array<unsigned long long, kSlotCount> returns_ {};
};

} // namespace vector_gpu
