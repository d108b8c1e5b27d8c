// host_copy.h -- host memory copies split over a small persistent pool of threads.
//
// Moving plan inputs and results between pageable host buffers and page-locked staging slots is a host
// memcpy; one thread moves ~10 GB/s, well below what a PCIe Gen5 x16 link carries, so the copy, not the
// DMA, bounded a host-column plan. The copy is split into contiguous parts done in parallel by the caller
// and the pool's workers (the pool is created on first use and lives until the process ends).
#pragma once

#include <cstddef>

namespace vgpu {

//! memcpy(dst, src, bytes), split across up to kHostCopyThreads threads for large copies. Thread-safe:
//! concurrent callers are served one after another.
void ParallelHostCopy(void *dst, const void *src, size_t bytes);

constexpr int kHostCopyThreads = 12;

} // namespace vgpu
