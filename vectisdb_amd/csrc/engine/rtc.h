// rtc.h -- hipRTC compilation of generated kernels for the current device (gfx950), with an on-disk cache
// of code objects so later processes skip the compiler.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace vgpu {

struct RtcKernel {
	hipModule_t module = nullptr;
	hipFunction_t function = nullptr;
	std::string name;
	bool from_disk = false;
	double compile_ms = 0;
};

//! Compiles `source` (one extern "C" kernel named `kernel_name`) with hipRTC for the current device and
//! loads it. `fast_math` swaps strict IEEE (-ffp-contract=off) for -ffast-math. Throws std::runtime_error
//! with the compiler log on failure. The returned object is owned by the caller (DestroyRtcKernel).
RtcKernel *CompileRtc(const std::string &source, const std::string &kernel_name, bool fast_math);
void DestroyRtcKernel(RtcKernel *kernel);

//! Architecture string used for compilation (gcnArchName of the current device; "gfx950" when none).
std::string DeviceArch();

//! The disk cache directory ("" = disabled): VECTOR_GPU_KERNEL_CACHE_DIR, else $XDG_CACHE_HOME or
//! ~/.cache + /vectis_hip_kernels. VECTOR_GPU_KERNEL_CACHE_DIR=off disables it.
std::string KernelDiskCacheDir();

} // namespace vgpu
