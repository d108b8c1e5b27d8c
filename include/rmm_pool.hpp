#pragma once
// rmm_pool.hpp -- device-memory budget and sub-allocating arena (drop-in for cuda_engine/include/rmm_pool.hpp).
//
// Production role, as in the reference: a headroom oracle. EnsureInitialized records a budget of a
// fraction of the device's currently free HBM (0.85 by default, VECTOR_GPU_VRAM_FRACTION overrides) and
// HasHeadroomFor answers routing's "does this plan's working set fit" question. The arena itself is only
// reserved by the first Allocate(), which the executor never calls (it draws from a stream-ordered
// caching pool). Allocate/Free sub-allocate that arena best-fit, 256-byte aligned, merging neighbouring
// holes on Free, with every live block's size tracked by the pool.

#include <cstddef>
#include <map>
#include <mutex>
#include <set>
#include <unordered_map>
#include <utility>

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

	//! Best-fit sub-allocation (the smallest hole that fits; the lowest address among equal holes);
	//! nullptr for 0 bytes, when uninitialized, when the arena cannot be reserved, or when no hole is large
	//! enough.
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

	void SetBudgetLocked(size_t bytes);
	bool ReserveArenaLocked();
	void AddHoleLocked(size_t offset, size_t size);
	void RemoveHoleLocked(size_t offset, size_t size);

	// Read by the inline TotalPoolBytes().
	mutable std::mutex mutex_;
	size_t total_pool_bytes_ = 0;

	bool initialized_ = false;
	void *arena_ = nullptr; // reserved by the first Allocate()
	// Free space as holes indexed twice: by offset (to merge neighbours on Free) and by (size, offset) (so
	// Allocate takes the smallest hole that fits and the headroom test reads the largest hole directly).
	std::map<size_t, size_t> holes_;
	std::set<std::pair<size_t, size_t>> holes_by_size_;
	std::unordered_map<size_t, size_t> taken_; // offset -> aligned size of each live allocation
	size_t taken_bytes_ = 0;
};

} // namespace vector_gpu
