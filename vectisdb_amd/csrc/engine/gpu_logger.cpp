// gpu_logger.cpp -- bounded two-file JSONL logger (contract of cuda_engine/include/gpu_logger.hpp:24-35).
#include "../../../include/gpu_logger.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <sstream>
#include <thread>

namespace vector_gpu {

namespace {

const char *EnvValue(const char *name) {
	const char *v = std::getenv(name);
	return (v != nullptr && v[0] != '\0') ? v : nullptr;
}

std::string LogFileName(const std::string &dir, int index) {
	std::string name = "gpu_log." + std::to_string(index) + ".jsonl";
	if (dir.empty()) {
		return name;
	}
	return dir.back() == '/' ? dir + name : dir + "/" + name;
}

const char *LevelText(LogLevel level) {
	switch (level) {
	case LogLevel::ERROR:
		return "error";
	case LogLevel::INFO:
		return "info";
	default:
		return "debug";
	}
}

} // namespace

GpuLogger &GpuLogger::Instance() {
	static GpuLogger *logger = new GpuLogger();
	return *logger;
}

GpuLogger::GpuLogger() {
	EnsureConfigured();
}

GpuLogger::~GpuLogger() {
	if (file_ != nullptr) {
		std::fclose(static_cast<std::FILE *>(file_));
	}
}

void GpuLogger::EnsureConfigured() {
	std::lock_guard<std::mutex> guard(mutex_);
	if (configured_) {
		return;
	}
	configured_ = true;
	const char *on = EnvValue("VECTOR_GPU_LOG");
	if (on == nullptr || std::strcmp(on, "0") == 0) {
		return;
	}
	if (const char *lvl = EnvValue("VECTOR_GPU_LOG_LEVEL")) {
		if (std::strcmp(lvl, "error") == 0) {
			level_.store(static_cast<int>(LogLevel::ERROR));
		} else if (std::strcmp(lvl, "debug") == 0) {
			level_.store(static_cast<int>(LogLevel::DEBUG));
		} else {
			level_.store(static_cast<int>(LogLevel::INFO));
		}
	}
	if (const char *cap = EnvValue("VECTOR_GPU_LOG_CAP_MIB")) {
		long mib = std::strtol(cap, nullptr, 10);
		if (mib > 0) {
			half_cap_bytes_ = static_cast<size_t>(mib) * 1024u * 1024u / 2u;
		}
	}
	if (const char *path = EnvValue("VECTOR_GPU_LOG_PATH")) {
		directory_ = path;
	}
	OpenActiveLocked(false);
	enabled_.store(file_ != nullptr, std::memory_order_relaxed);
}

void GpuLogger::OpenActiveLocked(bool truncate) {
	if (file_ != nullptr) {
		std::fclose(static_cast<std::FILE *>(file_));
		file_ = nullptr;
	}
	std::string name = LogFileName(directory_, active_file_);
	file_ = std::fopen(name.c_str(), truncate ? "wb" : "ab");
	active_bytes_ = 0;
	if (file_ != nullptr && !truncate) {
		std::fseek(static_cast<std::FILE *>(file_), 0, SEEK_END);
		long pos = std::ftell(static_cast<std::FILE *>(file_));
		active_bytes_ = pos > 0 ? static_cast<size_t>(pos) : 0;
	}
}

void GpuLogger::RotateIfNeededLocked(size_t incoming_bytes) {
	if (file_ == nullptr || active_bytes_ + incoming_bytes <= half_cap_bytes_) {
		return;
	}
	active_file_ = 1 - active_file_;
	OpenActiveLocked(true);
}

void GpuLogger::Log(LogLevel level, const char *event, const std::string &fields) noexcept {
	try {
		if (!Enabled(level)) {
			return;
		}
		auto now = std::chrono::duration_cast<std::chrono::microseconds>(
		               std::chrono::system_clock::now().time_since_epoch())
		               .count();
		std::ostringstream line;
		line << "{\"ts_us\":" << now << ",\"thread\":" << std::hash<std::thread::id> {}(std::this_thread::get_id())
		     << ",\"lvl\":\"" << LevelText(level) << "\",\"evt\":" << Quote(event != nullptr ? event : "?");
		if (!fields.empty()) {
			line << "," << fields;
		}
		line << "}\n";
		const std::string text = line.str();
		std::lock_guard<std::mutex> guard(mutex_);
		RotateIfNeededLocked(text.size());
		if (file_ == nullptr) {
			return;
		}
		std::fwrite(text.data(), 1, text.size(), static_cast<std::FILE *>(file_));
		std::fflush(static_cast<std::FILE *>(file_));
		active_bytes_ += text.size();
	} catch (...) {
	}
}

std::string GpuLogger::Quote(const std::string &value) {
	std::string out;
	out.reserve(value.size() + 2);
	out.push_back('"');
	for (unsigned char c : value) {
		switch (c) {
		case '"':
			out += "\\\"";
			break;
		case '\\':
			out += "\\\\";
			break;
		case '\n':
			out += "\\n";
			break;
		case '\r':
			out += "\\r";
			break;
		case '\t':
			out += "\\t";
			break;
		default:
			if (c < 0x20) {
				char buf[8];
				std::snprintf(buf, sizeof(buf), "\\u%04x", c);
				out += buf;
			} else {
				out.push_back(static_cast<char>(c));
			}
		}
	}
	out.push_back('"');
	return out;
}

void GpuLogger::ConfigureForTesting(bool enabled, LogLevel level, size_t cap_bytes, const std::string &directory) {
	std::lock_guard<std::mutex> guard(mutex_);
	configured_ = true;
	level_.store(static_cast<int>(level), std::memory_order_relaxed);
	half_cap_bytes_ = cap_bytes / 2u > 0 ? cap_bytes / 2u : 1u;
	directory_ = directory;
	active_file_ = 0;
	if (!enabled) {
		enabled_.store(false, std::memory_order_relaxed);
		if (file_ != nullptr) {
			std::fclose(static_cast<std::FILE *>(file_));
			file_ = nullptr;
		}
		return;
	}
	OpenActiveLocked(true);
	if (std::FILE *other = std::fopen(LogFileName(directory_, 1).c_str(), "wb")) {
		std::fclose(other);
	}
	enabled_.store(file_ != nullptr, std::memory_order_relaxed);
}

void GpuLogger::FlushForTesting() {
	std::lock_guard<std::mutex> guard(mutex_);
	if (file_ != nullptr) {
		std::fflush(static_cast<std::FILE *>(file_));
	}
}

size_t GpuLogger::OnDiskBytesForTesting() const {
	std::lock_guard<std::mutex> guard(mutex_);
	size_t total = 0;
	for (int i = 0; i < 2; i++) {
		if (std::FILE *f = std::fopen(LogFileName(directory_, i).c_str(), "rb")) {
			std::fseek(f, 0, SEEK_END);
			long size = std::ftell(f);
			total += size > 0 ? static_cast<size_t>(size) : 0;
			std::fclose(f);
		}
	}
	return total;
}

} // namespace vector_gpu
