// grouped_host.cpp -- grouped entry points (placeholder until the grouped
// kernel lands: every call reports NOTIMPLEMENTED_ERROR).
#include "api_internal.hpp"

extern "C" {
ssblasStatus_t ssblasGroupedPlanCreate(ssblasGroupedPlan_t* plan, SSINT, const SSINT[], const ssblasOperation_t[],
                                       const ssblasOperation_t[], const ssblasDataType_t[],
                                       const ssblasDataType_t[], const ssblasDataType_t[], const void* const[],
                                       const void* const[], const SSINT[], const SSINT[], const SSINT[],
                                       const void* const[], const SSINT[], const void* const[], const SSINT[],
                                       void* const[], const SSINT[], ssblasComputeType_t, ssblasGemmAlgo_t)
{
    if (plan) *plan = nullptr;
    return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;
}
ssblasStatus_t ssblasGroupedPlanExecute(ssblasGroupedPlan_t, ssblasStream_t) { return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR; }
ssblasStatus_t ssblasGroupedPlanGraphLaunch(ssblasGroupedPlan_t, ssblasStream_t) { return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR; }
ssblasStatus_t ssblasGroupedPlanDestroy(ssblasGroupedPlan_t) { return SSBLAS_STATUS_SUCCESS; }
ssblasStatus_t ssblasGemmGroupedBatchedEx(SSINT, const SSINT[], const ssblasOperation_t[], const ssblasOperation_t[],
                                          const ssblasDataType_t[], const ssblasDataType_t[],
                                          const ssblasDataType_t[], const void* const[], const void* const[],
                                          const SSINT[], const SSINT[], const SSINT[], const void* const[],
                                          const SSINT[], const void* const[], const SSINT[], void* const[],
                                          const SSINT[], ssblasComputeType_t, ssblasGemmAlgo_t)
{
    return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;
}
}
