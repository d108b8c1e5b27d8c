// device_pool.h -- stream-aware caching device allocator sized for one MI355X (288 GB HBM3E).
//
// hipMalloc/hipFree serialize the device, so the engine never calls them per operator: every plan
// temporary comes from this pool and goes back to it. A freed block is reusable at once on the stream
// that freed it (stream order protects it) and on any other stream once the event recorded at free time
// has completed. Large blocks are rounded to 2 MiB so repeated shapes (the same query over the same
// tables) find exact-size blocks again.
#pragma once

#include "../hip/kernels.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace vgpu {

class DevicePool {
public:
	static DevicePool &Instance();

	//! Throws HipError when the device is out of memory even after releasing the cache.
	void *Alloc(size_t bytes, hipStream_t stream);
	//! Returns a block to the cache; `stream` is the last stream that used it.
	void Free(void *ptr, hipStream_t stream);
	//! Returns several blocks last used on `stream` with one reuse event recorded for all of them (a plan's
	//! scratch arena frees dozens of blocks at once; one event per block cost tens of microseconds of host
	//! time per plan).
	void FreeBatch(const std::vector<void *> &ptrs, hipStream_t stream);
	//! Frees memory of unknown origin: a pool block goes back to the pool, its reuse event recorded on the
	//! stream it was allocated on (the null stream once that stream has been forgotten); anything else is
	//! released with hipFree. A caller that used the block on any other stream must have synchronized that
	//! stream first.
	void FreeAny(void *ptr);
	bool Owns(void *ptr) const;

	//! Releases every cached (free) block back to the driver. Synchronizes the device.
	void Trim();
	//! Call before destroying `stream`: synchronizes it and makes every cached block last used on it
	//! reusable on any stream, so no event recorded on it is touched after it is gone; live blocks allocated
	//! on it are re-tagged with the null stream.
	void ForgetStream(hipStream_t stream);

	size_t BytesInUse() const;
	size_t BytesCached() const;
	size_t PeakBytesInUse() const;
	void ResetPeak();

private:
	DevicePool() = default;
	~DevicePool();
	DevicePool(const DevicePool &) = delete;
	DevicePool &operator=(const DevicePool &) = delete;

	//! One event recorded when a batch of blocks was freed, shared by the blocks of the batch; destroyed with
	//! the last of them. Once seen complete it is not queried again.
	struct ReadyEvent {
		hipEvent_t event = nullptr;
		bool done = false;
		~ReadyEvent() {
			if (event != nullptr) {
				(void)hipEventDestroy(event);
			}
		}
	};
	struct Block {
		void *ptr = nullptr;
		size_t size = 0;
		hipStream_t stream = nullptr;
		std::shared_ptr<ReadyEvent> ready; // recorded at free; block reusable elsewhere once complete
		bool idle = false;                 // its stream is gone and was drained: reusable anywhere
	};
	static std::shared_ptr<ReadyEvent> RecordReady(hipStream_t stream);
	void CacheLocked(void *ptr, hipStream_t stream, const std::shared_ptr<ReadyEvent> &ready);
	static size_t RoundSize(size_t bytes);
	void *TakeCachedLocked(size_t size, hipStream_t stream);

	mutable std::mutex mutex_;
	std::multimap<size_t, Block> free_;             // by size
	struct Live {
		size_t size;        // rounded size
		hipStream_t stream; // stream the block was allocated on
	};
	std::unordered_map<void *, Live> live_;
	size_t in_use_ = 0;
	size_t cached_ = 0;
	size_t peak_ = 0;
};

//! Scratch arena over the pool: everything allocated is returned (tagged with `stream`) on destruction.
class PoolArena : public Scratch {
public:
	explicit PoolArena(hipStream_t stream) : stream_(stream) {
	}
	~PoolArena() override;
	PoolArena(const PoolArena &) = delete;
	PoolArena &operator=(const PoolArena &) = delete;

	void *Alloc(size_t bytes) override;
	//! Hands ownership of `ptr` (allocated by this arena) to the caller; it is no longer freed here.
	void Release(void *ptr);
	//! Takes ownership of `ptr`, a block obtained from DevicePool::Alloc outside this arena; it is returned
	//! to the pool with the arena's other blocks.
	void Adopt(void *ptr);
	hipStream_t stream() const {
		return stream_;
	}

private:
	hipStream_t stream_;
	std::vector<void *> blocks_;
};

} // namespace vgpu
