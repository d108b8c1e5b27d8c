#pragma once
// gpu_memory_pool.hpp -- fixed ring of page-locked batch slots feeding the stream pipeline (drop-in for
// cuda_engine/include/gpu_memory_pool.hpp).
//
// kSlotCount slots of kSlotBytes each are allocated with hipHostMalloc(hipHostMallocPortable) the first
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

	//! Allocates every slot on first use; throws std::runtime_error (freeing what was allocated) when one
	//! cannot be page-locked, and the next call tries again.
	static GpuMemoryPool &Instance();

	GpuMemoryPool(const GpuMemoryPool &) = delete;
	GpuMemoryPool &operator=(const GpuMemoryPool &) = delete;

	//! Blocks until the next slot in cyclic order is free, checks it out and advances the cursor.
	GpuMemoryPoolSlot AcquireNext();

	//! Test-only: rewinds the cursor to slot 0 (no slot may be checked out).
	void ResetCursorForTesting();

private:
	GpuMemoryPool();
	~GpuMemoryPool();
	friend class GpuMemoryPoolSlot;
	void Release(int index);

	struct Slot {
		void *data = nullptr;
		bool in_use = false;
	};
	std::mutex mutex_;
	std::condition_variable cv_;
	std::array<Slot, kSlotCount> slots_ {};
	size_t next_index_ = 0;
};

} // namespace vector_gpu
