// gpu_logger.cpp -- the engine's bounded JSONL activity log (the interface and on-disk format of
// cuda_engine/include/gpu_logger.hpp: gpu_log.0.jsonl / gpu_log.1.jsonl, one JSON object per line).
//
// The sink writes through POSIX file descriptors opened with O_APPEND, one write(2) per record, so a record is
// never split or interleaved and nothing sits in a user-space buffer (a crash loses no logged line). Two
// generations share the byte cap: when the active file would pass half of it, the other file is truncated
// and becomes the active one.
#include "../../../include/gpu_logger.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace vector_gpu {

struct GpuLogger::Sink {
	std::mutex mutex;
	std::string directory;
	size_t generation_cap = 25u << 20; // bytes per file: half the total cap
	int active = 0;                    // 0 or 1: which gpu_log.<n>.jsonl receives records
	int fd = -1;
	size_t written = 0; // bytes in the active file

	std::string PathOf(int generation) const {
		std::string name = "gpu_log." + std::to_string(generation) + ".jsonl";
		if (directory.empty()) {
			return name;
		}
		return directory.back() == '/' ? directory + name : directory + "/" + name;
	}
	void Close() {
		if (fd >= 0) {
			::close(fd);
			fd = -1;
		}
	}
	//! Makes `generation` the active file; `fresh` truncates it, else appends after what it holds.
	bool Open(int generation, bool fresh) {
		Close();
		active = generation;
		fd = ::open(PathOf(generation).c_str(), O_WRONLY | O_CREAT | O_APPEND | O_CLOEXEC | (fresh ? O_TRUNC : 0),
		            0644);
		written = 0;
		struct stat st;
		if (fd >= 0 && ::fstat(fd, &st) == 0) {
			written = static_cast<size_t>(st.st_size);
		}
		return fd >= 0;
	}
	void Append(const std::string &line) {
		if (written + line.size() > generation_cap && written > 0) {
			Open(1 - active, true);
		}
		if (fd < 0) {
			return;
		}
		const char *p = line.data();
		size_t left = line.size();
		while (left > 0) {
			const ssize_t n = ::write(fd, p, left);
			if (n <= 0) {
				return;
			}
			p += n;
			left -= static_cast<size_t>(n);
		}
		written += line.size();
	}
};

namespace {

const char *Env(const char *name) {
	const char *v = std::getenv(name);
	return v != nullptr && *v != '\0' ? v : nullptr; // set-but-empty counts as unset
}

LogLevel ParseLevel(const char *text) {
	if (std::strcmp(text, "error") == 0) {
		return LogLevel::ERROR;
	}
	return std::strcmp(text, "debug") == 0 ? LogLevel::DEBUG : LogLevel::INFO;
}

const char *LevelName(LogLevel level) {
	static const char *const kNames[] = {"error", "info", "debug"};
	const int i = static_cast<int>(level);
	return kNames[i < 0 ? 0 : (i > 2 ? 2 : i)];
}

} // namespace

GpuLogger &GpuLogger::Instance() {
	static GpuLogger *logger = new GpuLogger(); // never destroyed: call sites may log during static teardown
	return *logger;
}

GpuLogger::GpuLogger() : sink_(new Sink()) {
	// Configured here, not lazily on first Log(): VGPU_LOG tests Enabled() before calling Log().
	const char *on = Env("VECTOR_GPU_LOG");
	if (on == nullptr || std::strcmp(on, "0") == 0) {
		return;
	}
	if (const char *lvl = Env("VECTOR_GPU_LOG_LEVEL")) {
		level_.store(static_cast<int>(ParseLevel(lvl)), std::memory_order_relaxed);
	}
	if (const char *cap = Env("VECTOR_GPU_LOG_CAP_MIB")) {
		const long mib = std::strtol(cap, nullptr, 10);
		if (mib > 0) {
			sink_->generation_cap = (static_cast<size_t>(mib) << 20) / 2;
		}
	}
	if (const char *dir = Env("VECTOR_GPU_LOG_PATH")) {
		sink_->directory = dir;
	}
	enabled_.store(sink_->Open(0, false), std::memory_order_relaxed);
}

GpuLogger::~GpuLogger() {
	sink_->Close();
	delete sink_;
}

void GpuLogger::Log(LogLevel level, const char *event, const std::string &fields) noexcept {
	if (!Enabled(level)) {
		return;
	}
	try {
		const long long us = std::chrono::duration_cast<std::chrono::microseconds>(
		                         std::chrono::system_clock::now().time_since_epoch())
		                         .count();
		const long tid = static_cast<long>(::syscall(SYS_gettid));
		std::string line;
		line.reserve(96 + fields.size());
		line += "{\"ts_us\":" + std::to_string(us) + ",\"thread\":" + std::to_string(tid) + ",\"lvl\":\"";
		line += LevelName(level);
		line += "\",\"evt\":" + Quote(event != nullptr ? event : "?");
		if (!fields.empty()) {
			line += ',';
			line += fields;
		}
		line += "}\n";
		std::lock_guard<std::mutex> lock(sink_->mutex);
		sink_->Append(line);
	} catch (...) {
		// logging never fails the caller
	}
}

std::string GpuLogger::Quote(const std::string &value) {
	// Escape of each byte below 0x20 and of '"' / '\\'; other bytes pass through (UTF-8 stays as is).
	static const char *const kShort[32] = {
	    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, "\\b", "\\t", "\\n",
	    nullptr, "\\f",   "\\r",   nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
	    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
	std::string out = "\"";
	out.reserve(value.size() + 2);
	for (const char ch : value) {
		const unsigned char c = static_cast<unsigned char>(ch);
		if (c == '"' || c == '\\') {
			out += '\\';
			out += ch;
		} else if (c >= 0x20) {
			out += ch;
		} else if (kShort[c] != nullptr) {
			out += kShort[c];
		} else {
			char hex[8];
			std::snprintf(hex, sizeof(hex), "\\u%04x", c);
			out += hex;
		}
	}
	out += '"';
	return out;
}

void GpuLogger::ConfigureForTesting(bool enabled, LogLevel level, size_t cap_bytes, const std::string &directory) {
	std::lock_guard<std::mutex> lock(sink_->mutex);
	enabled_.store(false, std::memory_order_relaxed);
	level_.store(static_cast<int>(level), std::memory_order_relaxed);
	sink_->Close();
	sink_->directory = directory;
	sink_->generation_cap = cap_bytes / 2 > 0 ? cap_bytes / 2 : 1;
	if (!enabled) {
		return;
	}
	// both generations start empty
	const int other = ::open(sink_->PathOf(1).c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
	if (other >= 0) {
		::close(other);
	}
	enabled_.store(sink_->Open(0, true), std::memory_order_relaxed);
}

void GpuLogger::FlushForTesting() {
	// Records are written with write(2) as they are logged: nothing is buffered in this process.
	std::lock_guard<std::mutex> lock(sink_->mutex);
}

size_t GpuLogger::OnDiskBytesForTesting() const {
	std::lock_guard<std::mutex> lock(sink_->mutex);
	size_t total = 0;
	for (int g = 0; g < 2; g++) {
		struct stat st;
		if (::stat(sink_->PathOf(g).c_str(), &st) == 0) {
			total += static_cast<size_t>(st.st_size);
		}
	}
	return total;
}

} // namespace vector_gpu
