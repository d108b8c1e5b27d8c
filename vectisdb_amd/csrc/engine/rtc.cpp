// rtc.cpp -- hipRTC compile + code-object disk cache. Replaces the reference's NVRTC + PTX disk cache +
// cuModuleLoadDataEx path (cuda_engine/src/code_generator.cpp:108-212, :327-370): the artifact cached here
// is a gfx950 code object, loaded with hipModuleLoadData, keyed by SHA-256 of (source, options, arch,
// hipRTC version).
#include "rtc.h"

#include "../../../include/gpu_logger.hpp"
#include "../../../include/kernel_cache.hpp"

#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <unistd.h>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <vector>

namespace vgpu {

namespace {

std::vector<std::string> RtcOptions(bool fast_math, const std::string &arch) {
	std::vector<std::string> opts = {"--offload-arch=" + arch, "-O3", "-std=c++17"};
	if (fast_math) {
		opts.push_back("-ffast-math");
	} else {
		opts.push_back("-ffp-contract=off");
	}
	// Packed FP32 arithmetic stays off in every kernel this engine builds.
	opts.insert(opts.end(), {"-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"});
	return opts;
}

bool ReadFile(const std::string &path, std::string &out) {
	std::ifstream in(path, std::ios::binary);
	if (!in) {
		return false;
	}
	std::ostringstream ss;
	ss << in.rdbuf();
	out = ss.str();
	return !out.empty();
}

void WriteFileAtomic(const std::string &path, const std::string &data) {
	std::error_code ec;
	std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
	std::string tmp = path + ".tmp" + std::to_string(::getpid());
	{
		std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
		if (!out) {
			return;
		}
		out.write(data.data(), (std::streamsize)data.size());
		if (!out) {
			return;
		}
	}
	std::filesystem::rename(tmp, path, ec);
	if (ec) {
		std::filesystem::remove(tmp, ec);
	}
}

RtcKernel *LoadCodeObject(const std::string &code, const std::string &kernel_name) {
	auto kernel = std::make_unique<RtcKernel>();
	kernel->name = kernel_name;
	hipError_t st = hipModuleLoadData(&kernel->module, code.data());
	if (st != hipSuccess) {
		throw std::runtime_error(std::string("hipModuleLoadData failed: ") + hipGetErrorString(st));
	}
	st = hipModuleGetFunction(&kernel->function, kernel->module, kernel_name.c_str());
	if (st != hipSuccess) {
		(void)hipModuleUnload(kernel->module);
		throw std::runtime_error(std::string("hipModuleGetFunction failed: ") + hipGetErrorString(st));
	}
	return kernel.release();
}

} // namespace

std::string DeviceArch() {
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) {
		return "gfx950";
	}
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
		return "gfx950";
	}
	std::string arch = prop.gcnArchName;
	return arch.empty() ? "gfx950" : arch;
}

std::string KernelDiskCacheDir() {
	if (const char *dir = std::getenv("VECTOR_GPU_KERNEL_CACHE_DIR")) {
		if (std::strcmp(dir, "off") == 0 || dir[0] == '\0') {
			return "";
		}
		return dir;
	}
	if (const char *xdg = std::getenv("XDG_CACHE_HOME")) {
		if (xdg[0] != '\0') {
			return std::string(xdg) + "/vectis_hip_kernels";
		}
	}
	if (const char *home = std::getenv("HOME")) {
		if (home[0] != '\0') {
			return std::string(home) + "/.cache/vectis_hip_kernels";
		}
	}
	return "";
}

RtcKernel *CompileRtc(const std::string &source, const std::string &kernel_name, bool fast_math) {
	auto start = std::chrono::steady_clock::now();
	const std::string arch = DeviceArch();
	auto opts = RtcOptions(fast_math, arch);
	int major = 0, minor = 0;
	(void)hiprtcVersion(&major, &minor);
	std::string key_text = source + "\n// opts:";
	for (auto &o : opts) {
		key_text += " " + o;
	}
	key_text += "\n// hiprtc " + std::to_string(major) + "." + std::to_string(minor);
	const std::string key = vector_gpu::KernelCache::HashSource(key_text);
	const std::string dir = KernelDiskCacheDir();
	std::string path;
	if (!dir.empty()) {
		std::string safe_arch = arch;
		for (auto &c : safe_arch) {
			if (c == ':' || c == '/') {
				c = '_';
			}
		}
		path = dir + "/" + key + "." + safe_arch + ".hsaco";
		std::string cached;
		if (ReadFile(path, cached)) {
			try {
				RtcKernel *k = LoadCodeObject(cached, kernel_name);
				k->from_disk = true;
				k->compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
				return k;
			} catch (const std::exception &) {
				// stale or corrupt: rebuild below and overwrite
			}
		}
	}

	hiprtcProgram prog;
	if (hiprtcCreateProgram(&prog, source.c_str(), "vectis_fused.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
		throw std::runtime_error("hiprtcCreateProgram failed");
	}
	std::vector<const char *> copts;
	for (auto &o : opts) {
		copts.push_back(o.c_str());
	}
	hiprtcResult res = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
	size_t log_size = 0;
	(void)hiprtcGetProgramLogSize(prog, &log_size);
	std::string log;
	if (log_size > 1) {
		log.resize(log_size);
		(void)hiprtcGetProgramLog(prog, &log[0]);
	}
	if (res != HIPRTC_SUCCESS) {
		(void)hiprtcDestroyProgram(&prog);
		throw std::runtime_error("hipRTC compilation failed:\n" + log + "\n--- generated source ---\n" + source);
	}
	size_t code_size = 0;
	(void)hiprtcGetCodeSize(prog, &code_size);
	std::string code(code_size, '\0');
	(void)hiprtcGetCode(prog, &code[0]);
	(void)hiprtcDestroyProgram(&prog);
	RtcKernel *k = LoadCodeObject(code, kernel_name);
	if (!path.empty()) {
		WriteFileAtomic(path, code);
	}
	k->compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
	VGPU_LOG(vector_gpu::LogLevel::INFO, "compile",
	         "\"kernel\":" + vector_gpu::GpuLogger::Quote(kernel_name) + ",\"disk\":" +
	             vector_gpu::GpuLogger::Quote(path.empty() ? "off" : "miss") + ",\"ms\":" +
	             std::to_string(k->compile_ms));
	return k;
}

void DestroyRtcKernel(RtcKernel *kernel) {
	if (kernel == nullptr) {
		return;
	}
	if (kernel->module != nullptr) {
		(void)hipModuleUnload(kernel->module);
	}
	delete kernel;
}

} // namespace vgpu
