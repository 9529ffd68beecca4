// kernels_bf16x3.hip -- kernel instantiations of the bf16x3 arithmetic form.
#include "launch_impl.hpp"

namespace ssb {
hipError_t launch_bf16x3(const GemmProblem& g, int cfg, hipStream_t stream)
{
    return launch_form<ArithBF16X3>(g, cfg, stream);
}
}  // namespace ssb
