// onesweep_i32n.hip -- instantiation unit of the onesweep radix pass: 32 items per thread, payload loaded at its scatter (8-bit digits).
#include "onesweep_kernel.h"

namespace vgpu {
namespace sortpass {

template void LaunchVariant<8, 32, false, 256>(bool, bool, const KeyPack &, const void *, void *, const PayloadArgs &,
                                         uint64_t, int, int, const uint32_t *, uint32_t *, uint32_t *, hipStream_t,
                                         unsigned, const uint64_t *);

} // namespace sortpass
} // namespace vgpu
