// onesweep_i16p.hip -- instantiation unit of the onesweep radix pass: 16 items per thread, payload prefetched (8- and 11-bit digits).
#include "onesweep_kernel.h"

namespace vgpu {
namespace sortpass {

template void LaunchVariant<8, 16, true, 256>(bool, bool, const KeyPack &, const void *, void *, const PayloadArgs &,
                                         uint64_t, int, int, const uint32_t *, uint32_t *, uint32_t *, hipStream_t,
                                         unsigned, const uint64_t *);
template void LaunchVariant<11, 16, true, 256>(bool, bool, const KeyPack &, const void *, void *, const PayloadArgs &,
                                         uint64_t, int, int, const uint32_t *, uint32_t *, uint32_t *, hipStream_t,
                                         unsigned, const uint64_t *);

} // namespace sortpass
} // namespace vgpu
