// grouped_f64.hip -- grouped kernel of the family (ArithF64, ArithF64).
#include "grouped_kernel.hpp"

namespace ssb {
hipError_t launch_grouped_f64(const GroupedDevice& d, hipStream_t st)
{
    return run_grouped<ArithF64, ArithF64>(d, st);
}
}  // namespace ssb
