#pragma once
// code_generator.hpp -- per-expression kernels compiled with hipRTC for gfx950 (drop-in for
// cuda_engine/include/code_generator.hpp).
//
// Takes a GpuExpr body (a statement reading ((const T*)inputs[k])[row] and writing ((U*)out)[idx]), wraps
// it in a grid-stride kernel that resolves row through an optional uint32 selection vector and ANDs the
// referenced inputs' dense byte validity into out_valid, compiles it with hipRTC (strict IEEE by default,
// VECTOR_GPU_FAST_MATH=1 or SetGpuFastMathEnabled(true) for fast math), caches the loaded module in the
// KernelCache keyed by SHA-256(body + option tag) and on disk keyed by SHA-256 of the wrapped source.
//
// The plan executor does not launch through this class: it fuses every FILTER/PROJECTION expression of
// a pipeline into one generated kernel. CodeGenerator keeps the reference's single-expression contract
// for its other callers (the streaming pipeline, tests, tools).
//
// Launch passes the input and validity pointer tables to the kernel by value (kernel arguments), so it
// needs no device scratch, no host-to-device copy and no lock; LaunchPreloaded takes tables the caller
// already placed in device memory. Both are asynchronous on `stream` (a hipStream_t, nullptr = the null
// stream).

#include "gpu_engine.hpp"
#include "kernel_cache.hpp"

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace vector_gpu {

//! Fast-math selection for hipRTC: SetGpuFastMathEnabled overrides VECTOR_GPU_FAST_MATH ("1"/"true"/"TRUE").
bool IsGpuFastMathEnabled();
void SetGpuFastMathEnabled(bool enabled);

class CodeGenerator {
public:
	//! Registers the module deleter on `cache`, which then owns every handle inserted into it.
	explicit CodeGenerator(KernelCache &cache);
	~CodeGenerator();

	CodeGenerator(const CodeGenerator &) = delete;
	CodeGenerator &operator=(const CodeGenerator &) = delete;

	//! Returns the kernel for `expr`, compiling on a miss. The key is derived from the body text (the
	//! caller's source_hash is not trusted). The handle is borrowed from the cache: valid until the next
	//! insertion. Throws std::runtime_error carrying the hipRTC log on a compile failure, and
	//! std::logic_error when the cache declines to store it (zero capacity).
	CompiledKernelHandle CompileOrFetch(const GpuExpr &expr);

	//! Launches over `n` > 0 rows (std::invalid_argument otherwise, and for a null handle or an
	//! `input_valid` whose size is neither 0 nor inputs.size()). At most 16 inputs.
	void Launch(CompiledKernelHandle handle, const std::vector<const void *> &inputs, void *output, size_t n,
	            const uint32_t *selection = nullptr, void *stream = nullptr,
	            const std::vector<const bool *> &input_valid = {}, bool *out_valid = nullptr) const;

	//! Same, with the pointer tables already in device memory (`inputs_device`: `input_count` device
	//! pointers; `input_valid_device`: nullptr or `input_count` entries).
	void LaunchPreloaded(CompiledKernelHandle handle, const void *const *inputs_device, int input_count,
	                     void *output, size_t n, const uint32_t *selection = nullptr, void *stream = nullptr,
	                     const bool *const *input_valid_device = nullptr, bool *out_valid = nullptr) const;

	//! The kernel source generated around `expression_body` (exposed for tests).
	std::string WrapGridStrideLoop(const std::string &expression_body) const;

private:
	KernelCache &cache_;
	std::mutex compile_mutex_;
};

} // namespace vector_gpu
