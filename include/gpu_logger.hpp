#pragma once
// gpu_logger.hpp -- bounded JSONL activity log (drop-in for cuda_engine/include/gpu_logger.hpp).
//
// Off by default; a disabled call site costs one relaxed atomic load (Enabled is inline). Records go to
// gpu_log.0.jsonl / gpu_log.1.jsonl, each capped at half the total cap; when the active file would
// overflow, the other (older) file is truncated and becomes active, so the newest records always survive.
// Environment (read once): VECTOR_GPU_LOG (0/1), VECTOR_GPU_LOG_LEVEL (error|info|debug),
// VECTOR_GPU_LOG_CAP_MIB (default 50), VECTOR_GPU_LOG_PATH (directory).

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

namespace vector_gpu {

enum class LogLevel : int {
	ERROR = 0,
	INFO = 1,
	DEBUG = 2,
};

class GpuLogger {
public:
	static GpuLogger &Instance();

	bool Enabled(LogLevel level) const {
		return enabled_.load(std::memory_order_relaxed) &&
		       static_cast<int>(level) <= level_.load(std::memory_order_relaxed);
	}

	//! One line {"ts_us":..,"thread":..,"lvl":"..","evt":"<event>",<fields>}; `fields` is a JSON object body
	//! without braces. Thread-safe, never throws, flushed per record.
	void Log(LogLevel level, const char *event, const std::string &fields) noexcept;

	//! JSON string literal (quotes included) with escapes.
	static std::string Quote(const std::string &value);

	void ConfigureForTesting(bool enabled, LogLevel level, size_t cap_bytes, const std::string &directory);
	void FlushForTesting();
	size_t OnDiskBytesForTesting() const;

private:
	GpuLogger();
	~GpuLogger();
	GpuLogger(const GpuLogger &) = delete;
	GpuLogger &operator=(const GpuLogger &) = delete;

	// Read by the inline Enabled() on every VGPU_LOG: the only state callers touch.
	std::atomic<bool> enabled_ {false};
	std::atomic<int> level_ {static_cast<int>(LogLevel::INFO)};
	// Everything else (directory, files, byte counts) lives in the sink, behind its own mutex.
	struct Sink;
	Sink *sink_ = nullptr;
};

} // namespace vector_gpu

#define VGPU_LOG(level, event, fields_expr)                                                                      \
	do {                                                                                                         \
		auto &_vgpu_logger = ::vector_gpu::GpuLogger::Instance();                                                \
		if (_vgpu_logger.Enabled(level)) {                                                                       \
			_vgpu_logger.Log((level), (event), (fields_expr));                                                   \
		}                                                                                                        \
	} while (0)
