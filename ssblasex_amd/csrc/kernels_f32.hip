// kernels_f32.hip -- kernel instantiations of the f32 arithmetic form.
#include "launch_impl.hpp"

namespace ssb {
hipError_t launch_f32(const GemmProblem& g, int cfg, hipStream_t stream)
{
    return launch_form<ArithF32>(g, cfg, stream);
}
}  // namespace ssb
