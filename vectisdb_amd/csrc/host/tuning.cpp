// tuning.cpp -- the sort autotuner's decision table.
//
// Picks the digit width (and therefore the number of onesweep passes) from the packed key width and the
// row count. The table is measured, not derived: `bench.py --sweep` times every (radix_bits, passes)
// candidate per shape on the device and the winners are recorded here. VGPU_RADIX_BITS overrides it.
#include "operators.h"

#include <cstdlib>

namespace vgpu {

namespace {
//! Optional overrides for sweeps: VGPU_SORT_ITEMS (16/32), VGPU_SORT_PREFETCH (0/1), VGPU_SORT_BLOCK
//! (256/512/1024; 512 and 1024 run with 16 items and no prefetch only).
void ApplyVariantOverrides(SortTuning &t) {
	if (const char *env = std::getenv("VGPU_SORT_BLOCK")) {
		const int b = std::atoi(env);
		t.block = b >= 1024 ? 1024 : b >= 512 ? 512 : 256;
		if (t.block >= 512) {
			t.items = 16;
			t.prefetch = 0;
		}
		t.fixed = true;
	}
	if (const char *env = std::getenv("VGPU_SORT_ITEMS")) {
		t.items = std::atoi(env) >= 32 ? 32 : 16;
		t.fixed = true;
	}
	if (const char *env = std::getenv("VGPU_SORT_PREFETCH")) {
		t.prefetch = std::atoi(env) != 0 ? 1 : 0;
		t.fixed = true;
	}
}
} // namespace

SortTuning TuneSort(int total_bits, uint64_t rows) {
	SortTuning t;
	ApplyVariantOverrides(t);
	if (const char *env = std::getenv("VGPU_RADIX_BITS")) {
		int v = std::atoi(env);
		t.radix_bits = v >= 11 ? 11 : 8;
		t.fixed = true;
		return t;
	}
	// Measured on MI355X, 100M rows, 20-bit keys + 16 B payload: three 8-bit passes take 4.8 ms, two 11-bit
	// passes 10.0 ms. An 11-bit digit leaves ~2 items per bucket in a 4096-item tile, so the scatter writes
	// 8-byte fragments and the look-back walks 2048 counters per tile; 8-bit digits keep 16-item runs.
	// Only inputs too small to fill the device (launch-bound, < 64K rows) take the wider digit to save a
	// pass.
	int passes8 = (total_bits + 7) / 8;
	int passes11 = (total_bits + 10) / 11;
	t.radix_bits = (rows < (1ull << 16) && passes11 < passes8) ? 11 : 8;
	return t;
}

} // namespace vgpu
