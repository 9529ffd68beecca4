// grouped_i8_bf16x3.hip -- grouped kernel of the family (ArithI8, ArithBF16X3).
#include "grouped_kernel.hpp"

namespace ssb {
hipError_t launch_grouped_i8_bf16x3(const GroupedDevice& d, hipStream_t st)
{
    return run_grouped<ArithI8, ArithBF16X3>(d, st);
}
}  // namespace ssb
