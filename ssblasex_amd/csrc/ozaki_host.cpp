// ozaki_host.cpp -- placeholder until the Ozaki kernels land.
#include "api_internal.hpp"

namespace ssb {
hipError_t launch_ozaki(const GemmProblem&, int, hipStream_t, ssblasStatus_t* st)
{
    *st = SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;
    return hipSuccess;
}
}  // namespace ssb
