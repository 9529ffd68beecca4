// kernels_i8.hip -- kernel instantiations of the i8 arithmetic form.
#include "launch_impl.hpp"

namespace ssb {
hipError_t launch_i8(const GemmProblem& g, int cfg, hipStream_t stream)
{
    return launch_form<ArithI8>(g, cfg, stream);
}
}  // namespace ssb
