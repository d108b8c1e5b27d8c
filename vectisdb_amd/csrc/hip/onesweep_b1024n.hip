// onesweep_b1024n.hip -- instantiation unit of the onesweep radix pass: 1024 threads x 16 items (16384-row
// tile), payload loaded at its scatter (8-bit digits).
#include "onesweep_kernel.h"

namespace vgpu {
namespace sortpass {

template void LaunchVariant<8, 16, false, 1024>(bool, bool, const KeyPack &, const void *, void *,
                                                const PayloadArgs &, uint64_t, int, int, const uint32_t *,
                                                uint32_t *, uint32_t *, hipStream_t, unsigned, const uint64_t *);

} // namespace sortpass
} // namespace vgpu
