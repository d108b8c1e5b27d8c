#pragma once
// kernel_cache.hpp -- LRU of compiled fused kernels keyed by SHA-256 hex of their source
// (drop-in for cuda_engine/include/kernel_cache.hpp).

#include <cstddef>
#include <list>
#include <mutex>
#include <string>
#include <unordered_map>

namespace vector_gpu {

//! Opaque compiled-kernel handle (a loaded hipModule + function inside the engine).
using CompiledKernelHandle = void *;

class KernelCache {
public:
	//! Frees a handle the cache owns (on eviction, replacement, decline or destruction).
	using HandleDeleter = void (*)(CompiledKernelHandle);

	explicit KernelCache(size_t max_entries = 256) : max_entries_(max_entries) {
	}
	~KernelCache();

	KernelCache(const KernelCache &) = delete;
	KernelCache &operator=(const KernelCache &) = delete;

	//! Registers the owner's deleter; registering a different one later throws std::logic_error.
	void SetHandleDeleter(HandleDeleter deleter);

	//! Returns the handle (and marks it most recently used) or nullptr.
	CompiledKernelHandle Get(const std::string &source_hash);

	//! Stores `handle`, evicting the least recently used entry at capacity. A zero-capacity cache stores
	//! nothing and frees the declined handle.
	void Insert(const std::string &source_hash, CompiledKernelHandle handle);

	//! Lowercase hex SHA-256 (FIPS 180-4) of `cuda_source`.
	static std::string HashSource(const std::string &cuda_source);

private:
	void DeleteHandleLocked(CompiledKernelHandle handle);

	std::mutex mutex_;
	HandleDeleter deleter_ = nullptr;
	size_t max_entries_;
	std::list<std::string> lru_order_; // front = most recently used
	std::unordered_map<std::string, std::pair<CompiledKernelHandle, std::list<std::string>::iterator>> entries_;
};

} // namespace vector_gpu
