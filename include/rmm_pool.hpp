#pragma once
// rmm_pool.hpp -- device-memory budget and sub-allocating arena (drop-in for cuda_engine/include/rmm_pool.hpp).
//
// Production role, as in the reference: a headroom oracle. EnsureInitialized records a budget of a
// fraction of the device's currently free HBM (0.85 by default, VECTOR_GPU_VRAM_FRACTION overrides) and
// HasHeadroomFor answers routing's "does this plan's working set fit" question. The arena itself is only
// reserved by the first Allocate(), which the executor never calls (it draws from a stream-ordered
// caching pool); Allocate/Free keep the reference's first-fit, coalescing, size-tracked semantics so
// tests and any future caller see the same behaviour.

#include <cstddef>
#include <map>
#include <mutex>
#include <unordered_map>

namespace vector_gpu {

class RmmPool {
public:
	static RmmPool &Instance();

	//! Records a budget of `fraction_of_vram` (0, 1] of currently free device memory; no-op once
	//! initialized. VECTOR_GPU_VRAM_FRACTION in (0, 1] overrides the argument. Throws
	//! std::invalid_argument for a fraction outside (0, 1], std::runtime_error when the device cannot be
	//! queried or the budget rounds to 0 bytes.
	void EnsureInitialized(double fraction_of_vram = 0.85);

	//! True when a single free block of the (256-byte aligned) size exists. False, never throwing, when
	//! uninitialized, when `estimated_bytes` is 0, or when it exceeds the whole budget (which also keeps
	//! the alignment arithmetic away from the top of size_t).
	bool HasHeadroomFor(size_t estimated_bytes) const;

	//! First-fit sub-allocation; nullptr for 0 bytes, when uninitialized, when the arena cannot be
	//! reserved, or when no free block is large enough.
	void *Allocate(size_t bytes);

	//! Returns a block from Allocate(). The tracked size is used (`bytes` is accepted for interface
	//! compatibility); nullptr, foreign, interior and already-freed pointers are ignored.
	void Free(void *ptr, size_t bytes);

	size_t TotalPoolBytes() const {
		std::lock_guard<std::mutex> guard(mutex_);
		return total_pool_bytes_;
	}

	size_t AllocatedBytes() const;

	//! Test-only: releases the arena and forgets the budget.
	void ResetForTesting();

	//! Test-only: initializes to exactly `bytes` (rounded up to 256) and reserves the arena now; no-op
	//! once initialized.
	void EnsureInitializedExactBytesForTesting(size_t bytes);

private:
	RmmPool();
	~RmmPool();
	RmmPool(const RmmPool &) = delete;
	RmmPool &operator=(const RmmPool &) = delete;

	void AllocateArenaLocked(size_t requested_bytes);
	bool EnsureArenaLocked();
	void ResetBlocksLocked(size_t total);

	mutable std::mutex mutex_;
	bool initialized_ = false;
	void *arena_ = nullptr;
	size_t total_pool_bytes_ = 0;
	std::map<size_t, size_t> free_blocks_;           // offset -> size, coalesced
	std::unordered_map<size_t, size_t> live_blocks_; // offset -> aligned size
};

} // namespace vector_gpu
