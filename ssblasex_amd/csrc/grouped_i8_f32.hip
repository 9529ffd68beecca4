// grouped_i8_f32.hip -- grouped kernel of the family (ArithI8, ArithF32).
#include "grouped_kernel.hpp"

namespace ssb {
hipError_t launch_grouped_i8_f32(const GroupedDevice& d, hipStream_t st)
{
    return run_grouped<ArithI8, ArithF32>(d, st);
}
}  // namespace ssb
