// code_generator.cpp -- single-expression kernels through hipRTC (contract of
// cuda_engine/src/code_generator.cpp:269-561: grid-stride wrapper, row = sel[idx] or idx, out_valid = AND
// of the referenced inputs' validity, strict IEEE unless fast math is selected, cache key
// SHA-256(body + " [opt:strict]" | " [opt:fast_math]"), borrowed handles, zero-capacity caches refused).
//
// The generated module has two entry points around one device function holding the body: one takes the
// input/validity pointer tables by value in its kernel arguments (Launch: no device scratch, no copy, no
// lock, safe from any number of streams), the other reads caller-owned device tables (LaunchPreloaded).
#include "../../../include/code_generator.hpp"
#include "../../../include/gpu_logger.hpp"
#include "rtc.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>

namespace vector_gpu {

namespace {

constexpr int kMaxInputs = 16;
constexpr unsigned kBlock = 256;

struct ExprKernel {
	vgpu::RtcKernel *rtc = nullptr;
	hipFunction_t by_value = nullptr;
	hipFunction_t preloaded = nullptr;
};

void DestroyExprKernel(CompiledKernelHandle handle) {
	auto *k = static_cast<ExprKernel *>(handle);
	if (k != nullptr) {
		vgpu::DestroyRtcKernel(k->rtc);
		delete k;
	}
}

std::atomic<int> g_fast_math_override {-1};

struct TableArgs {
	const void *inputs[kMaxInputs];
	const bool *valid[kMaxInputs];
};

unsigned GridFor(size_t n) {
	size_t g = (n + kBlock - 1) / kBlock;
	return static_cast<unsigned>(g < 8192 ? (g == 0 ? 1 : g) : 8192);
}

} // namespace

bool IsGpuFastMathEnabled() {
	const int v = g_fast_math_override.load();
	if (v != -1) {
		return v != 0;
	}
	const char *env = std::getenv("VECTOR_GPU_FAST_MATH");
	return env != nullptr && (std::strcmp(env, "1") == 0 || std::strcmp(env, "true") == 0 ||
	                          std::strcmp(env, "TRUE") == 0);
}

void SetGpuFastMathEnabled(bool enabled) {
	g_fast_math_override.store(enabled ? 1 : 0);
}

CodeGenerator::CodeGenerator(KernelCache &cache) : cache_(cache) {
	cache_.SetHandleDeleter(&DestroyExprKernel);
}

CodeGenerator::~CodeGenerator() = default;

std::string CodeGenerator::WrapGridStrideLoop(const std::string &expression_body) const {
	std::ostringstream o;
	o << "__device__ __forceinline__ double __vgpu_radians(double d) { return d * 0.017453292519943295; }\n"
	     "__device__ __forceinline__ double __vgpu_degrees(double r) { return r * 57.29577951308232; }\n"
	     "struct __vgpu_tables { const void* inputs["
	  << kMaxInputs << "]; const bool* valid[" << kMaxInputs
	  << "]; };\n"
	     "__device__ __forceinline__ void __vgpu_eval(const void* const* inputs, void* out, unsigned long long idx,\n"
	     "                                            unsigned long long row) {\n"
	     "  (void)inputs; (void)out; (void)idx; (void)row;\n"
	     "  "
	  << expression_body
	  << "\n"
	     "}\n"
	     "__device__ __forceinline__ void __vgpu_loop(const void* const* inputs, const bool* const* valid, int count,\n"
	     "                                            void* out, unsigned long long n, const unsigned int* sel,\n"
	     "                                            bool* out_valid) {\n"
	     "  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;\n"
	     "  for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;\n"
	     "       idx += stride) {\n"
	     "    const unsigned long long row = sel == 0 ? idx : (unsigned long long)sel[idx];\n"
	     "    if (out_valid != 0) {\n"
	     "      bool ok = true;\n"
	     "      if (valid != 0) {\n"
	     "        for (int k = 0; k < count; k++) {\n"
	     "          if (valid[k] != 0 && !valid[k][row]) ok = false;\n"
	     "        }\n"
	     "      }\n"
	     "      out_valid[idx] = ok;\n"
	     "    }\n"
	     "    __vgpu_eval(inputs, out, idx, row);\n"
	     "  }\n"
	     "}\n"
	     "extern \"C\" __global__ void __launch_bounds__("
	  << kBlock
	  << ") vectis_expr_tab(__vgpu_tables t, int count, int has_valid, void* out,\n"
	     "    unsigned long long n, const unsigned int* sel, bool* out_valid) {\n"
	     "  const void* inputs["
	  << kMaxInputs << "]; const bool* valid[" << kMaxInputs
	  << "];\n"
	     "#pragma unroll\n"
	     "  for (int k = 0; k < "
	  << kMaxInputs
	  << "; k++) { inputs[k] = t.inputs[k]; valid[k] = t.valid[k]; }\n"
	     "  __vgpu_loop(inputs, has_valid ? valid : 0, count, out, n, sel, out_valid);\n"
	     "}\n"
	     "extern \"C\" __global__ void __launch_bounds__("
	  << kBlock
	  << ") vectis_expr_dev(const void* const* inputs, const bool* const* valid,\n"
	     "    int count, void* out, unsigned long long n, const unsigned int* sel, bool* out_valid) {\n"
	     "  __vgpu_loop(inputs, valid, count, out, n, sel, out_valid);\n"
	     "}\n";
	return o.str();
}

CompiledKernelHandle CodeGenerator::CompileOrFetch(const GpuExpr &expr) {
	const bool fast_math = IsGpuFastMathEnabled();
	const std::string key =
	    KernelCache::HashSource(expr.generated_cuda_source + (fast_math ? " [opt:fast_math]" : " [opt:strict]"));
	std::lock_guard<std::mutex> guard(compile_mutex_);
	const auto start = std::chrono::steady_clock::now();
	if (auto cached = cache_.Get(key)) {
		VGPU_LOG(LogLevel::DEBUG, "compile", "\"hit\":true,\"key\":" + GpuLogger::Quote(key.substr(0, 16)));
		return cached;
	}
	const std::string source = WrapGridStrideLoop(expr.generated_cuda_source);
	auto *k = new ExprKernel();
	try {
		k->rtc = vgpu::CompileRtc(source, "vectis_expr_tab", fast_math);
		k->by_value = k->rtc->function;
		hipError_t st = hipModuleGetFunction(&k->preloaded, k->rtc->module, "vectis_expr_dev");
		if (st != hipSuccess) {
			throw std::runtime_error(std::string("CodeGenerator: hipModuleGetFunction failed: ") +
			                         hipGetErrorString(st));
		}
	} catch (...) {
		DestroyExprKernel(k);
		throw;
	}
	VGPU_LOG(LogLevel::INFO, "compile",
	         "\"hit\":false,\"disk\":" + GpuLogger::Quote(k->rtc->from_disk ? "hit" : "miss") + ",\"key\":" +
	             GpuLogger::Quote(key.substr(0, 16)) + ",\"us\":" +
	             std::to_string(std::chrono::duration_cast<std::chrono::microseconds>(
	                                std::chrono::steady_clock::now() - start)
	                                .count()));
	cache_.Insert(key, k);
	// Re-fetch: a zero-capacity cache declines (and frees) the insert.
	auto stored = cache_.Get(key);
	if (stored == nullptr) {
		throw std::logic_error("CodeGenerator::CompileOrFetch: the kernel cache declined to store the compiled "
		                       "kernel (zero-capacity cache?) -- CodeGenerator requires a cache with capacity >= 1");
	}
	return stored;
}

void CodeGenerator::Launch(CompiledKernelHandle handle, const std::vector<const void *> &inputs, void *output,
                           size_t n, const uint32_t *selection, void *stream,
                           const std::vector<const bool *> &input_valid, bool *out_valid) const {
	if (n == 0) {
		throw std::invalid_argument("CodeGenerator::Launch: n must be > 0 (caller should special-case empty input "
		                            "rather than launching a zero-element grid)");
	}
	if (handle == nullptr) {
		throw std::invalid_argument("CodeGenerator::Launch: handle must not be null -- pass the value returned by "
		                            "CompileOrFetch");
	}
	if (!input_valid.empty() && input_valid.size() != inputs.size()) {
		throw std::invalid_argument("CodeGenerator::Launch: input_valid must be empty or exactly inputs.size() "
		                            "entries (one slot per referenced column, nullptr where that column has no "
		                            "nulls)");
	}
	if (inputs.size() > static_cast<size_t>(kMaxInputs)) {
		throw std::invalid_argument("CodeGenerator::Launch: at most 16 inputs (use LaunchPreloaded for more)");
	}
	auto *k = static_cast<ExprKernel *>(handle);
	TableArgs tables {};
	for (size_t i = 0; i < inputs.size(); i++) {
		tables.inputs[i] = inputs[i];
		tables.valid[i] = input_valid.empty() ? nullptr : input_valid[i];
	}
	int count = static_cast<int>(inputs.size());
	int has_valid = input_valid.empty() ? 0 : 1;
	unsigned long long rows = n;
	void *args[] = {&tables, &count, &has_valid, &output, &rows, &selection, &out_valid};
	hipError_t st = hipModuleLaunchKernel(k->by_value, GridFor(n), 1, 1, kBlock, 1, 1, 0,
	                                      static_cast<hipStream_t>(stream), args, nullptr);
	if (st != hipSuccess) {
		throw std::runtime_error(std::string("CodeGenerator::Launch: hipModuleLaunchKernel failed: ") +
		                         hipGetErrorString(st));
	}
}

void CodeGenerator::LaunchPreloaded(CompiledKernelHandle handle, const void *const *inputs_device, int input_count,
                                    void *output, size_t n, const uint32_t *selection, void *stream,
                                    const bool *const *input_valid_device, bool *out_valid) const {
	if (n == 0) {
		throw std::invalid_argument("CodeGenerator::LaunchPreloaded: n must be > 0 (caller should special-case "
		                            "empty input rather than launching a zero-element grid)");
	}
	if (handle == nullptr) {
		throw std::invalid_argument("CodeGenerator::LaunchPreloaded: handle must not be null -- pass the value "
		                            "returned by CompileOrFetch");
	}
	if (input_count < 0 || (input_count > 0 && inputs_device == nullptr)) {
		throw std::invalid_argument("CodeGenerator::LaunchPreloaded: input_count must be >= 0 and inputs_device "
		                            "must be non-null whenever it is > 0");
	}
	auto *k = static_cast<ExprKernel *>(handle);
	unsigned long long rows = n;
	void *args[] = {&inputs_device, &input_valid_device, &input_count, &output, &rows, &selection, &out_valid};
	hipError_t st = hipModuleLaunchKernel(k->preloaded, GridFor(n), 1, 1, kBlock, 1, 1, 0,
	                                      static_cast<hipStream_t>(stream), args, nullptr);
	if (st != hipSuccess) {
		throw std::runtime_error(std::string("CodeGenerator::LaunchPreloaded: hipModuleLaunchKernel failed: ") +
		                         hipGetErrorString(st));
	}
}

} // namespace vector_gpu
