#include "device_pool.h"

#include <algorithm>
#include <string>

namespace vgpu {

namespace {
constexpr size_t kSmallGranule = 256;
constexpr size_t kLargeGranule = 2u << 20;
constexpr size_t kLargeThreshold = 1u << 20;
} // namespace

DevicePool &DevicePool::Instance() {
	// Leaked on purpose: destroying it at exit would race the HIP runtime's own teardown.
	static DevicePool *pool = new DevicePool();
	return *pool;
}

DevicePool::~DevicePool() = default;

size_t DevicePool::RoundSize(size_t bytes) {
	if (bytes == 0) {
		bytes = 1;
	}
	if (bytes < kLargeThreshold) {
		return (bytes + kSmallGranule - 1) / kSmallGranule * kSmallGranule;
	}
	return (bytes + kLargeGranule - 1) / kLargeGranule * kLargeGranule;
}

void *DevicePool::TakeCachedLocked(size_t size, hipStream_t stream) {
	// Accept blocks up to 2x (small) or size + 25% (large) so a huge block is not burned on a tiny request.
	size_t limit = size < kLargeThreshold ? size * 2 : size + size / 4 + kLargeGranule;
	for (auto it = free_.lower_bound(size); it != free_.end() && it->first <= limit; ++it) {
		Block &b = it->second;
		bool usable = b.idle || b.stream == stream;
		if (!usable && b.ready != nullptr) {
			if (!b.ready->done) {
				b.ready->done = hipEventQuery(b.ready->event) == hipSuccess;
				if (!b.ready->done) {
					(void)hipGetLastError(); // "not ready" is an answer, not an error to report later
				}
			}
			usable = b.ready->done;
		}
		if (!usable) {
			continue;
		}
		void *ptr = b.ptr;
		size_t bsize = b.size;
		free_.erase(it); // drops this block's share of its event (an event is never recorded twice)
		cached_ -= bsize;
		live_[ptr] = Live {bsize, stream};
		in_use_ += bsize;
		peak_ = std::max(peak_, in_use_);
		return ptr;
	}
	return nullptr;
}

void *DevicePool::Alloc(size_t bytes, hipStream_t stream) {
	const size_t size = RoundSize(bytes);
	{
		std::lock_guard<std::mutex> lock(mutex_);
		if (void *p = TakeCachedLocked(size, stream)) {
			return p;
		}
	}
	void *ptr = nullptr;
	hipError_t st = hipMalloc(&ptr, size);
	if (st != hipSuccess) {
		(void)hipGetLastError();
		Trim();
		st = hipMalloc(&ptr, size);
		if (st != hipSuccess) {
			(void)hipGetLastError();
			throw HipError("device pool: hipMalloc(" + std::to_string(size) +
			               ") failed: " + hipGetErrorString(st));
		}
	}
	std::lock_guard<std::mutex> lock(mutex_);
	live_[ptr] = Live {size, stream};
	in_use_ += size;
	peak_ = std::max(peak_, in_use_);
	return ptr;
}

std::shared_ptr<DevicePool::ReadyEvent> DevicePool::RecordReady(hipStream_t stream) {
	auto ready = std::make_shared<ReadyEvent>();
	if (hipEventCreateWithFlags(&ready->event, hipEventDisableTiming) != hipSuccess) {
		(void)hipGetLastError();
		ready->event = nullptr;
		return nullptr;
	}
	if (hipEventRecord(ready->event, stream) != hipSuccess) {
		(void)hipGetLastError();
		return nullptr; // the holder destroys the event
	}
	return ready;
}

void DevicePool::CacheLocked(void *ptr, hipStream_t stream, const std::shared_ptr<ReadyEvent> &ready) {
	auto it = live_.find(ptr);
	if (it == live_.end()) {
		return; // not ours; never free foreign memory
	}
	Block b;
	b.ptr = ptr;
	b.size = it->second.size;
	b.stream = stream;
	b.ready = ready;
	live_.erase(it);
	in_use_ -= b.size;
	cached_ += b.size;
	free_.emplace(b.size, b);
}

void DevicePool::Free(void *ptr, hipStream_t stream) {
	if (ptr == nullptr) {
		return;
	}
	{
		std::lock_guard<std::mutex> lock(mutex_);
		if (live_.count(ptr) == 0) {
			return;
		}
	}
	const auto ready = RecordReady(stream);
	std::lock_guard<std::mutex> lock(mutex_);
	CacheLocked(ptr, stream, ready);
}

void DevicePool::FreeBatch(const std::vector<void *> &ptrs, hipStream_t stream) {
	if (ptrs.empty()) {
		return;
	}
	const auto ready = RecordReady(stream);
	std::lock_guard<std::mutex> lock(mutex_);
	for (void *p : ptrs) {
		if (p != nullptr) {
			CacheLocked(p, stream, ready);
		}
	}
}

bool DevicePool::Owns(void *ptr) const {
	std::lock_guard<std::mutex> lock(mutex_);
	return live_.count(ptr) != 0;
}

void DevicePool::FreeAny(void *ptr) {
	if (ptr == nullptr) {
		return;
	}
	hipStream_t last = nullptr;
	bool ours = false;
	{
		std::lock_guard<std::mutex> lock(mutex_);
		auto it = live_.find(ptr);
		if (it != live_.end()) {
			ours = true;
			last = it->second.stream;
		}
	}
	if (ours) {
		// The stream the block was allocated on is taken as its last user: the reuse event is recorded there.
		Free(ptr, last);
		return;
	}
	(void)hipFree(ptr);
}

void DevicePool::ForgetStream(hipStream_t stream) {
	(void)hipStreamSynchronize(stream);
	std::lock_guard<std::mutex> lock(mutex_);
	for (auto &kv : live_) {
		if (kv.second.stream == stream) {
			kv.second.stream = nullptr; // a later FreeAny records on the null stream, never on a dead handle
		}
	}
	for (auto &kv : free_) {
		Block &b = kv.second;
		if (b.stream != stream || b.idle) {
			continue;
		}
		b.ready.reset();
		b.stream = nullptr;
		b.idle = true;
	}
}

void DevicePool::Trim() {
	(void)hipDeviceSynchronize();
	std::lock_guard<std::mutex> lock(mutex_);
	for (auto &kv : free_) {
		(void)hipFree(kv.second.ptr);
	}
	free_.clear(); // and with the blocks, their events
	cached_ = 0;
}

size_t DevicePool::BytesInUse() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return in_use_;
}

size_t DevicePool::BytesCached() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return cached_;
}

size_t DevicePool::PeakBytesInUse() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return peak_;
}

void DevicePool::ResetPeak() {
	std::lock_guard<std::mutex> lock(mutex_);
	peak_ = in_use_;
}

PoolArena::~PoolArena() {
	DevicePool::Instance().FreeBatch(blocks_, stream_);
}

void *PoolArena::Alloc(size_t bytes) {
	void *p = DevicePool::Instance().Alloc(bytes, stream_);
	blocks_.push_back(p);
	return p;
}

void PoolArena::Adopt(void *ptr) {
	if (ptr != nullptr) {
		blocks_.push_back(ptr);
	}
}

void PoolArena::Release(void *ptr) {
	auto it = std::find(blocks_.begin(), blocks_.end(), ptr);
	if (it != blocks_.end()) {
		blocks_.erase(it);
	}
}

} // namespace vgpu
