// kernels_f64.hip -- kernel instantiations of the f64 arithmetic form.
#include "launch_impl.hpp"

namespace ssb {
hipError_t launch_f64(const GemmProblem& g, int cfg, hipStream_t stream)
{
    return launch_form<ArithF64>(g, cfg, stream);
}
}  // namespace ssb
