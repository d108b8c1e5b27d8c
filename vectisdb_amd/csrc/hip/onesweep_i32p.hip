// onesweep_i32p.hip -- instantiation unit of the onesweep radix pass: 32 items per thread, payload prefetched (8-bit digits).
#include "onesweep_kernel.h"

namespace vgpu {
namespace sortpass {

template void LaunchVariant<8, 32, true, 256>(bool, bool, const KeyPack &, const void *, void *, const PayloadArgs &,
                                         uint64_t, int, int, const uint32_t *, uint32_t *, uint32_t *, hipStream_t,
                                         unsigned, const uint64_t *);

} // namespace sortpass
} // namespace vgpu
