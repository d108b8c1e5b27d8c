#pragma once
// pinned_buffer_pool.hpp -- process-wide cache of page-locked host staging buffers (engine-private, like
// cuda_engine/include/pinned_buffer_pool.hpp; only the executor's upload/download staging uses it).
//
// Buffers come in power-of-two size classes (1 MiB and up) and are page-locked from birth with
// hipHostMalloc, so DMA engines read them directly (no registration of existing memory is attempted).
// Acquire hands out the smallest idle buffer whose class covers the request; a returned buffer goes back to
// its class's idle stack (most recently used first). At most kMaxSlots buffers exist at once: at the cap an
// idle buffer of a too-small class is released to make room, and with none idle Acquire blocks until a
// buffer comes back.

#include <condition_variable>
#include <cstddef>
#include <map>
#include <mutex>
#include <vector>

namespace vector_gpu {

//! One checked-out buffer. Move-only; goes back to the pool on destruction.
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
	//! The buffer's size class (>= the request).
	size_t bytes() const {
		return bytes_;
	}

private:
	friend class PinnedBufferPool;
	PinnedBufferSlot(class PinnedBufferPool *pool, void *data, size_t bytes) : pool_(pool), data_(data), bytes_(bytes) {
	}
	class PinnedBufferPool *pool_ = nullptr;
	void *data_ = nullptr;
	size_t bytes_ = 0;
};

class PinnedBufferPool {
public:
	static constexpr size_t kMaxSlots = 16;
	static constexpr size_t kMinClassBytes = 1u << 20;

	static PinnedBufferPool &Instance();

	//! A buffer of at least `bytes` (rounded up to its size class). Blocks while kMaxSlots buffers are checked
	//! out. Throws std::runtime_error when the page-locked allocation fails.
	PinnedBufferSlot Acquire(size_t bytes);

	//! Buffers that exist (checked out or idle) and the bytes they hold.
	size_t LiveBuffers() const;
	size_t LiveBytes() const;

private:
	PinnedBufferPool() = default;
	~PinnedBufferPool();
	PinnedBufferPool(const PinnedBufferPool &) = delete;
	PinnedBufferPool &operator=(const PinnedBufferPool &) = delete;

	friend class PinnedBufferSlot;
	void Return(void *data, size_t bytes);

	mutable std::mutex mutex_;
	std::condition_variable returned_;
	std::map<size_t, std::vector<void *>> idle_; // size class -> idle buffers, most recently returned last
	size_t live_buffers_ = 0;
	size_t live_bytes_ = 0;
};

} // namespace vector_gpu
