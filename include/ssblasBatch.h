/*
 * ssblasBatch.h -- C ABI of the MI355X (gfx950) implementation of
 * ssblasGemmBatchedEx.
 *
 * Drop-in for include/ssblasBatch.h of hirokitokura/ssblasEx: the first two
 * declarations below have exactly the reference's parameter lists and are
 * exported from libssblasGemmBatchedEx.so under the same names, so a caller
 * linked with -lssblasGemmBatchedEx against the reference links against this
 * library unchanged.  Everything after them is appended.
 *
 *   C[i] = alpha * op(A[i]) * op(B[i]) + beta * C[i],   0 <= i < batchCount
 *
 * op(A[i]) is m x k, op(B[i]) is k x n, C[i] is m x n, all column-major with
 * leading dimensions lda / ldb / ldc shared by every instance.
 *
 * Supported (Atype, Btype, Ctype) and the type alpha/beta are read as:
 *   (R_32F, R_32F, R_32F)  float
 *   (R_64F, R_64F, R_64F)  double
 *   (R_8I,  R_8I,  R_32I)  int32  (the reference code reads int32; its
 *                                   README table says R_8I -- code wins)
 *
 * Pointer arrays: A, B and C must be device-accessible arrays (device
 * memory, managed memory, or pinned host memory) of device pointers.  A
 * pageable host array is rejected with SSBLAS_STATUS_INVALID_VALUE.  alpha
 * and beta are host pointers read once by value.
 *
 * Status codes (same values as the reference):
 *   SUCCESS          computed (or nothing to do: k == 0 leaves C unchanged)
 *   INVALID_VALUE    unsupported type triple; OP_C with any type; unknown op;
 *                    negative m/n/k/batchCount; ld smaller than the stored
 *                    row count; pageable host pointer array; computeType
 *                    naming a form that does not fit the type triple
 *   INTERNAL_ERROR   m == 0, n == 0 or batchCount == 0 (the reference
 *                    returns it for zero tasks); a HIP launch failure
 *   ALLOC_FAILED     workspace allocation failed
 */
#ifndef SSBLASBATCH
#define SSBLASBATCH

#include "ssblasBatch_type.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A HIP stream (hipStream_t has this same type). NULL = the null stream. */
typedef struct ihipStream_t* ssblasStream_t;

/*
 * Replaces the reference entry of the same name
 * (include/ssblasBatch.h:15-35; Impl/GemmBatchedEx/ssblasGemmBatchedEx.cpp:173-220).
 * Synchronous: the call returns after C has been written, as the
 * reference's does.  Work runs on the null stream.
 * SSBLAS_GEMMBATCHEDEX_DEBUG set in the environment prints the arguments
 * (reference ssblasGemmBatchedEx.cpp:48-93).
 */
ssblasStatus_t ssblasGemmBatchedEx(
    ssblasOperation_t    transa,
    ssblasOperation_t    transb,
    SSINT                m,
    SSINT                n,
    SSINT                k,
    const void*          alpha,
    const void* const    A[],
    ssblasDataType_t     Atype,
    const SSINT          lda,
    const void* const    B[],
    ssblasDataType_t     Btype,
    const SSINT          ldb,
    const void*          beta,
    void* const          C[],
    ssblasDataType_t     Ctype,
    const SSINT          ldc,
    const SSINT          batchCount,
    ssblasComputeType_t  computeType,
    ssblasGemmAlgo_t     algo);

/*
 * Replaces the reference's ssblasGemmBatchedEx_64 (include/ssblasBatch.h:37-57).
 * The reference body is a stub returning an uninitialised value
 * (ssblasGemmBatchedEx.cpp:222-268); here it is the same computation as
 * ssblasGemmBatchedEx with 64-bit sizes (SSLONG == SSINT == long).
 */
ssblasStatus_t ssblasGemmBatchedEx_64(
    ssblasOperation_t    transa,
    ssblasOperation_t    transb,
    SSLONG               m,
    SSLONG               n,
    SSLONG               k,
    const void*          alpha,
    const void* const    A[],
    ssblasDataType_t     Atype,
    const SSLONG         lda,
    const void* const    B[],
    ssblasDataType_t     Btype,
    const SSLONG         ldb,
    const void*          beta,
    void* const          C[],
    ssblasDataType_t     Ctype,
    const SSLONG         ldc,
    const SSLONG         batchCount,
    ssblasComputeType_t  computeType,
    ssblasGemmAlgo_t     algo);

/* ======================= appended entry points ========================== */

/*
 * Asynchronous form of ssblasGemmBatchedEx (no counterpart in the
 * reference): same arguments and status table, work is enqueued on `stream`
 * and the call returns without waiting.  Allocates nothing for the plain
 * fp32/fp64/int8 and BF16 forms, so it may be captured into a HIP graph;
 * the Ozaki form uses a library-owned workspace that grows on demand
 * (capture it only after one uncaptured call of the same or larger size).
 */
ssblasStatus_t ssblasGemmBatchedExAsync(
    ssblasOperation_t    transa,
    ssblasOperation_t    transb,
    SSINT                m,
    SSINT                n,
    SSINT                k,
    const void*          alpha,
    const void* const    A[],
    ssblasDataType_t     Atype,
    const SSINT          lda,
    const void* const    B[],
    ssblasDataType_t     Btype,
    const SSINT          ldb,
    const void*          beta,
    void* const          C[],
    ssblasDataType_t     Ctype,
    const SSINT          ldc,
    const SSINT          batchCount,
    ssblasComputeType_t  computeType,
    ssblasGemmAlgo_t     algo,
    ssblasStream_t       stream);

/*
 * ssblasGemmBatchedExAsync for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, status
 * table and results as ssblasGemmBatchedExAsync; it calls no hipMalloc or
 * hipFree.  The device workspace of the Ozaki form comes from a fixed
 * scratch array of SSBLAS_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a call that needs more returns SSBLAS_STATUS_ALLOC_FAILED
 * before any device work (the plain and BF16 forms and the fused Ozaki form
 * for m, n, k <= 32 need none).  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams, other threads) race
 * on it; calls on one stream are ordered.
 */
#define SSBLAS_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

ssblasStatus_t ssblasGemmBatchedExNoAlloc(
    ssblasOperation_t    transa,
    ssblasOperation_t    transb,
    SSINT                m,
    SSINT                n,
    SSINT                k,
    const void*          alpha,
    const void* const    A[],
    ssblasDataType_t     Atype,
    const SSINT          lda,
    const void* const    B[],
    ssblasDataType_t     Btype,
    const SSINT          ldb,
    const void*          beta,
    void* const          C[],
    ssblasDataType_t     Ctype,
    const SSINT          ldc,
    const SSINT          batchCount,
    ssblasComputeType_t  computeType,
    ssblasGemmAlgo_t     algo,
    ssblasStream_t       stream);

/*
 * Grouped (variable-size) batched GEMM (no counterpart in the reference).
 *
 * Instances are numbered 0 .. total-1, total = sum(group_size[g]); group g
 * holds the next group_size[g] instances.  Per GROUP (host arrays of length
 * group_count): transa/transb, Atype/Btype/Ctype (each group one of the
 * three supported triples; groups may differ), alpha/beta (host pointers
 * to scalars of the group's scale type).  Per INSTANCE (host arrays of
 * length total): m, n, k, lda, ldb, ldc.  A, B, C: device-accessible arrays
 * of `total` device pointers.  computeType applies to every fp32 group
 * (DEFAULT/32F or a BF16 form); fp64 groups use native fp64.
 *
 * A plan holds the device work list (one entry per output tile, sorted
 * longest first) built once from the sizes; executing it is ONE persistent
 * kernel launch that reads the pointer arrays at run time, so the same plan
 * can be re-run after the pointed-to data (or the pointers) change.
 */
typedef struct ssblasGroupedPlan* ssblasGroupedPlan_t;

ssblasStatus_t ssblasGroupedPlanCreate(
    ssblasGroupedPlan_t*     plan,
    SSINT                    group_count,
    const SSINT              group_size[],
    const ssblasOperation_t  transa_array[],
    const ssblasOperation_t  transb_array[],
    const ssblasDataType_t   Atype_array[],
    const ssblasDataType_t   Btype_array[],
    const ssblasDataType_t   Ctype_array[],
    const void* const        alpha_array[],
    const void* const        beta_array[],
    const SSINT              m_array[],
    const SSINT              n_array[],
    const SSINT              k_array[],
    const void* const        A[],
    const SSINT              lda_array[],
    const void* const        B[],
    const SSINT              ldb_array[],
    void* const              C[],
    const SSINT              ldc_array[],
    ssblasComputeType_t      computeType,
    ssblasGemmAlgo_t         algo);

/* Enqueue the plan's single persistent launch on `stream` (capturable).
 * Launches of one plan must not overlap in time: workgroups claim the plan's
 * work through counters held in the plan (zero again when a launch ends). */
ssblasStatus_t ssblasGroupedPlanExecute(ssblasGroupedPlan_t plan, ssblasStream_t stream);

/* First call captures ssblasGroupedPlanExecute into a HIP graph owned by the
 * plan; every call then launches that graph on `stream`. */
ssblasStatus_t ssblasGroupedPlanGraphLaunch(ssblasGroupedPlan_t plan, ssblasStream_t stream);

ssblasStatus_t ssblasGroupedPlanDestroy(ssblasGroupedPlan_t plan);

/* One-shot synchronous grouped call: create, execute on the null stream,
 * wait, destroy. */
ssblasStatus_t ssblasGemmGroupedBatchedEx(
    SSINT                    group_count,
    const SSINT              group_size[],
    const ssblasOperation_t  transa_array[],
    const ssblasOperation_t  transb_array[],
    const ssblasDataType_t   Atype_array[],
    const ssblasDataType_t   Btype_array[],
    const ssblasDataType_t   Ctype_array[],
    const void* const        alpha_array[],
    const void* const        beta_array[],
    const SSINT              m_array[],
    const SSINT              n_array[],
    const SSINT              k_array[],
    const void* const        A[],
    const SSINT              lda_array[],
    const void* const        B[],
    const SSINT              ldb_array[],
    void* const              C[],
    const SSINT              ldc_array[],
    ssblasComputeType_t      computeType,
    ssblasGemmAlgo_t         algo);

/*
 * Error report of the Ozaki fp64 form (no counterpart in the reference).
 *
 * Every call with SSBLAS_COMPUTE_64F_EMULATED_OZAKI[_S<n>] computes, per
 * instance, an a-priori bound on its normwise error
 *     ||C - C_exact||_F / (|alpha| ||op(A)||_F ||op(B)||_F)
 * from the slice digits and the row / column sums of the operands (the
 * truncation of the slices, the product triangle and the fp64 folds; the
 * final rounding to fp64 that every fp64 result carries comes on top).  With
 * the automatic compute type (..._OZAKI) each instance takes the smallest
 * slice count, 3 to 9, whose bound is at most `target` (2^-55, a quarter of
 * fp64's unit roundoff); an instance that misses it at 9 slices keeps 9 and
 * is counted in over_target.  Rows or columns holding inf / nan give a bound
 * of +inf (their outputs are NaN).
 *
 * ssblasGetOzakiReport waits for `stream` and returns the report of the last
 * Ozaki call enqueued on it from this device (NOT_COMPUTED_ERROR if none):
 * the summary, and per instance the bound and the slice count used (the
 * first `capacity` instances into bounds[] / slices[], either may be NULL).
 */
typedef struct {
    SSINT  instances;     /* batchCount of the call */
    int    slices_min;    /* fewest / most slices any instance used */
    int    slices_max;
    double bound_max;     /* worst instance's bound */
    double target;        /* 2^-55 */
    SSINT  over_target;   /* instances whose bound exceeds target */
} ssblasOzakiReport_t;

ssblasStatus_t ssblasGetOzakiReport(
    ssblasStream_t        stream,
    ssblasOzakiReport_t*  report,
    float*                bounds,
    signed char*          slices,
    SSINT                 capacity);

/*
 * Which kernel a call with these arguments would run (no counterpart in the
 * reference; for tests and tuning): writes the tile-config name (e.g.
 * "dma256k2", "tile256x128", "direct16") or the Ozaki kernel family
 * ("ozaki_small", "ozaki_mma") into name[0 .. name_len-1].  Same status
 * table as ssblasGemmBatchedEx for the arguments it takes; touches no
 * device memory.
 */
ssblasStatus_t ssblasGetKernelConfig(
    ssblasOperation_t    transa,
    ssblasOperation_t    transb,
    SSINT                m,
    SSINT                n,
    SSINT                k,
    ssblasDataType_t     Atype,
    ssblasDataType_t     Btype,
    ssblasDataType_t     Ctype,
    const SSINT          lda,
    const SSINT          ldb,
    const SSINT          ldc,
    const SSINT          batchCount,
    ssblasComputeType_t  computeType,
    char*                name,
    SSINT                name_len);

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int ssblasGetVersion(void);

#ifdef __cplusplus
}
#endif

#endif /* SSBLASBATCH */
