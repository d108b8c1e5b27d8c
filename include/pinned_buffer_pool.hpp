#pragma once
// pinned_buffer_pool.hpp -- process-wide pool of page-locked host staging slots (drop-in for
// cuda_engine/include/pinned_buffer_pool.hpp).
//
// Same classes, members and semantics as the reference: a slot is checked out per concurrently
// transferring buffer, grows to the largest request it has served and never shrinks, and returns itself
// to the pool when the PinnedBufferSlot is destroyed. The pool grows lazily up to kMaxSlots; beyond that
// Acquire blocks until a slot is released. Memory comes from hipHostMalloc (page-locked from birth, so
// DMA engines read it directly; no registration of existing memory is ever attempted).

#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <vector>

namespace vector_gpu {

//! One checked-out slot. Move-only; releases itself on destruction.
class PinnedBufferSlot {
public:
	PinnedBufferSlot() = default;
	PinnedBufferSlot(const PinnedBufferSlot &) = delete;
	PinnedBufferSlot &operator=(const PinnedBufferSlot &) = delete;
	PinnedBufferSlot(PinnedBufferSlot &&other) noexcept;
	PinnedBufferSlot &operator=(PinnedBufferSlot &&other) noexcept;
	~PinnedBufferSlot();

	//! Page-locked host memory of at least the bytes requested in Acquire().
	void *data() const {
		return data_;
	}

private:
	friend class PinnedBufferPool;
	PinnedBufferSlot(class PinnedBufferPool *pool, int index, void *data) : pool_(pool), index_(index), data_(data) {
	}
	class PinnedBufferPool *pool_ = nullptr;
	int index_ = -1;
	void *data_ = nullptr;
};

class PinnedBufferPool {
public:
	static constexpr size_t kMaxSlots = 16;

	static PinnedBufferPool &Instance();

	//! Returns a free slot of at least `bytes`, growing the pool (below kMaxSlots) or the slot's
	//! allocation as needed; blocks while every slot is checked out at the cap. Throws
	//! std::runtime_error when the page-locked allocation fails.
	PinnedBufferSlot Acquire(size_t bytes);

private:
	PinnedBufferPool();
	~PinnedBufferPool();
	PinnedBufferPool(const PinnedBufferPool &) = delete;
	PinnedBufferPool &operator=(const PinnedBufferPool &) = delete;

	friend class PinnedBufferSlot;
	void Release(int index);

	struct Slot {
		void *data = nullptr;
		size_t capacity = 0;
		bool in_use = false;
	};
	std::mutex mutex_;
	std::condition_variable cv_;
	std::vector<Slot> slots_;
};

} // namespace vector_gpu
