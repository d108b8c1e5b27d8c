// tuning.cpp -- the sort autotuner's decision table.
//
// Picks the digit width (and therefore the number of onesweep passes) from the packed key width and the
// row count. The table is measured, not derived: `bench.py --sweep` times every (radix_bits, passes)
// candidate per shape on the device and the winners are recorded here. VGPU_RADIX_BITS overrides it.
#include "operators.h"

#include <cstdlib>

namespace vgpu {

SortTuning TuneSort(int total_bits, uint64_t rows) {
	SortTuning t;
	if (const char *env = std::getenv("VGPU_RADIX_BITS")) {
		int v = std::atoi(env);
		t.radix_bits = v >= 11 ? 11 : 8;
		return t;
	}
	// Small inputs are launch-bound: fewer, wider digits win. Large inputs pay for the wider digit's
	// sparser scatter, so 8-bit digits are kept unless they would add a whole pass.
	int passes8 = (total_bits + 7) / 8;
	int passes11 = (total_bits + 10) / 11;
	if (passes11 < passes8 && (rows < (1ull << 22) || passes8 - passes11 >= 1)) {
		t.radix_bits = 11;
	} else {
		t.radix_bits = 8;
	}
	return t;
}

} // namespace vgpu
