// grouped_i8_bf16x6.hip -- grouped kernel of the family (ArithI8, ArithBF16X6).
#include "grouped_kernel.hpp"

namespace ssb {
hipError_t launch_grouped_i8_bf16x6(const GroupedDevice& d, hipStream_t st)
{
    return run_grouped<ArithI8, ArithBF16X6>(d, st);
}
}  // namespace ssb
