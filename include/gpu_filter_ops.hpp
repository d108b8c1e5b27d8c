#pragma once
// gpu_filter_ops.hpp -- the stand-alone compaction primitive (drop-in for cuda_engine/include/gpu_filter_ops.hpp).

#include <cstddef>
#include <cstdint>

namespace vector_gpu {

//! Copies the elements of `input` (device, `n` elements) that are > `threshold` into `output` (device,
//! room for `n`), in input order, and writes their count to `*out_count` (host) before returning.
//! One gfx950 kernel: per-tile wave64 ballot ranking plus a decoupled look-back for the tile offset, on
//! the null stream. n == 0 writes 0 and touches neither pointer. Throws std::invalid_argument for a null
//! out_count, null buffers with n > 0, or input == output (partial overlap stays the caller's contract),
//! and std::runtime_error when a HIP call fails. An error latched by an earlier, unrelated HIP call is
//! cleared on entry rather than attributed to this one.
void FilterGreaterThanInt32(const int32_t *input, size_t n, int32_t threshold, int32_t *output, size_t *out_count);

} // namespace vector_gpu
