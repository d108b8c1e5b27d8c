// host_copy.cpp -- see host_copy.h.
#include "host_copy.h"

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace vgpu {
namespace {

//! Below this size one thread copies (waking workers costs more than it saves).
constexpr size_t kParallelMin = 2u << 20;

class CopyPool {
public:
	CopyPool() {
		for (int i = 0; i < kHostCopyThreads - 1; i++) {
			workers_.emplace_back([this, i] { Run(i); });
		}
	}
	~CopyPool() {
		{
			std::lock_guard<std::mutex> lock(mu_);
			stop_ = true;
		}
		cv_.notify_all();
		for (auto &t : workers_) {
			t.join();
		}
	}

	void Copy(void *dst, const void *src, size_t bytes) {
		std::lock_guard<std::mutex> serial(call_mu_); // one copy job at a time
		const size_t parts = kHostCopyThreads;
		// parts of equal size rounded to 4 KiB, the last takes the rest
		const size_t part = ((bytes / parts) + 4095) & ~size_t(4095);
		{
			std::lock_guard<std::mutex> lock(mu_);
			dst_ = static_cast<uint8_t *>(dst);
			src_ = static_cast<const uint8_t *>(src);
			bytes_ = bytes;
			part_ = part;
			pending_ = kHostCopyThreads - 1;
			generation_++;
		}
		cv_.notify_all();
		CopyPart(kHostCopyThreads - 1); // the caller takes the last part
		std::unique_lock<std::mutex> lock(mu_);
		done_cv_.wait(lock, [this] { return pending_ == 0; });
	}

private:
	void CopyPart(int index) const {
		const size_t begin = std::min(bytes_, (size_t)index * part_);
		const size_t end = index == kHostCopyThreads - 1 ? bytes_ : std::min(bytes_, begin + part_);
		if (end > begin) {
			std::memcpy(dst_ + begin, src_ + begin, end - begin);
		}
	}
	void Run(int index) {
		uint64_t seen = 0;
		for (;;) {
			{
				std::unique_lock<std::mutex> lock(mu_);
				cv_.wait(lock, [&] { return stop_ || generation_ != seen; });
				if (stop_) {
					return;
				}
				seen = generation_;
			}
			CopyPart(index);
			{
				std::lock_guard<std::mutex> lock(mu_);
				pending_--;
			}
			done_cv_.notify_one();
		}
	}

	std::mutex call_mu_;
	std::mutex mu_;
	std::condition_variable cv_, done_cv_;
	std::vector<std::thread> workers_;
	uint8_t *dst_ = nullptr;
	const uint8_t *src_ = nullptr;
	size_t bytes_ = 0, part_ = 0;
	int pending_ = 0;
	uint64_t generation_ = 0;
	bool stop_ = false;
};

} // namespace

void ParallelHostCopy(void *dst, const void *src, size_t bytes) {
	if (bytes < kParallelMin) {
		std::memcpy(dst, src, bytes);
		return;
	}
	static CopyPool *pool = new CopyPool(); // lives until exit: workers must not be joined during teardown
	pool->Copy(dst, src, bytes);
}

} // namespace vgpu
