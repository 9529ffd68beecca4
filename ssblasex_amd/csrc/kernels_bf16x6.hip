// kernels_bf16x6.hip -- kernel instantiations of the bf16x6 arithmetic form.
#include "launch_impl.hpp"

namespace ssb {
hipError_t launch_bf16x6(const GemmProblem& g, int cfg, hipStream_t stream)
{
    return launch_form<ArithBF16X6>(g, cfg, stream);
}
}  // namespace ssb
