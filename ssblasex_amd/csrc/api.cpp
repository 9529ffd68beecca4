// api.cpp -- the extern "C" entry points of libssblasGemmBatchedEx.so.
//
// Host-only C++ (compiled with the host compiler): argument checks that
// mirror the reference's status table, the optional argument dump, the
// choice of arithmetic form and tile config, and the kernel launch through
// the internal launch interface.  Every check runs before the first HIP
// call, so the whole status table is testable on a machine without a GPU.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <climits>
#include <mutex>

#include "../../include/ssblasBatch.h"
#include "api_internal.hpp"
#include "launch.hpp"

namespace ssb {

// ---------------------------------------------------------------------------
// Names for the argument dump (format of the reference's
// ssblasGemmBatchedExShowdebug, Impl/GemmBatchedEx/ssblasGemmBatchedEx.cpp:8-93)
// ---------------------------------------------------------------------------
static const char* op_name(ssblasOperation_t op)
{
    switch (op) {
    case SSBLAS_OP_N: return "SSBLAS_OP_N";
    case SSBLAS_OP_T: return "SSBLAS_OP_T";
    case SSBLAS_OP_C: return "SSBLAS_OP_C";
    default: return "UNKNOWN";
    }
}

static const char* dtype_name(ssblasDataType_t t)
{
    switch (t) {
    case SSBLAS_R_8I: return "SSBLAS_R_8I";
    case SSBLAS_R_8U: return "SSBLAS_R_8U";
    case SSBLAS_R_16F: return "SSBLAS_R_16F";
    case SSBLAS_R_32I: return "SSBLAS_R_32I";
    case SSBLAS_R_32F: return "SSBLAS_R_32F";
    case SSBLAS_R_64F: return "SSBLAS_R_64F";
    default: return "UNKNONW";
    }
}

static const char* compute_name(ssblasComputeType_t c)
{
    switch ((int)c) {
    case SSBLAS_COMPUTE_DEFAULT_TYPE: return "SSBLAS_COMPUTE_DEFAULT_TYPE";
    case SSBLAS_COMPUTE_32I: return "SSBLAS_COMPUTE_32I";
    case SSBLAS_COMPUTE_32F: return "SSBLAS_COMPUTE_32F";
    case SSBLAS_COMPUTE_64F: return "SSBLAS_COMPUTE_64F";
    case SSBLAS_COMPUTE_32F_FAST_BF16X3: return "SSBLAS_COMPUTE_32F_FAST_BF16X3";
    case SSBLAS_COMPUTE_32F_EMULATED_BF16X6: return "SSBLAS_COMPUTE_32F_EMULATED_BF16X6";
    case SSBLAS_COMPUTE_64F_EMULATED_OZAKI: return "SSBLAS_COMPUTE_64F_EMULATED_OZAKI";
    default:
        if ((int)c >= SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S3 && (int)c <= SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S9)
            return "SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S<n>";
        return "UNKNONW";
    }
}

static void debug_dump(ssblasOperation_t transa, ssblasOperation_t transb, long m, long n, long k,
                       ssblasDataType_t Atype, long lda, ssblasDataType_t Btype, long ldb,
                       ssblasDataType_t Ctype, long ldc, long batchCount,
                       ssblasComputeType_t computeType, ssblasGemmAlgo_t algo)
{
    printf("ssblasGemmBatchedExShowdebug: START\n");
    printf("transa: %s\n", op_name(transa));
    printf("transb: %s\n", op_name(transb));
    printf("M: %ld\n", m);
    printf("N: %ld\n", n);
    printf("K: %ld\n", k);
    printf("alpha: \n");
    printf("A: \n");
    printf("Atype: %s\n", dtype_name(Atype));
    printf("lda: %ld\n", lda);
    printf("B: \n");
    printf("Btype: %s\n", dtype_name(Btype));
    printf("ldb: %ld\n", ldb);
    printf("beta: \n");
    printf("C: \n");
    printf("Ctype: %s\n", dtype_name(Ctype));
    printf("ldc: %ld\n", ldc);
    printf("batchCount: %ld\n", batchCount);
    printf("computeType: %s\n", compute_name(computeType));
    printf("algo: %s\n", algo == SSBLAS_COMPUTE_DEFAULT ? "SSBLAS_COMPUTE_DEFAULT" : "UNKNONW");
    printf("ssblasGemmBatchedExShowdebug: END\n");
    fflush(stdout);
}

// ---------------------------------------------------------------------------
// Argument checks (order and codes: see the header's status table)
// ---------------------------------------------------------------------------
TypeTriple classify_types(ssblasDataType_t Atype, ssblasDataType_t Btype, ssblasDataType_t Ctype)
{
    // Reference Impl/GemmBatchedEx/ssblasGemmBatchedExSelector.hpp:35-107.
    if (Atype == SSBLAS_R_8I && Btype == SSBLAS_R_8I && Ctype == SSBLAS_R_32I) return TT_I8;
    if (Atype == SSBLAS_R_64F && Btype == SSBLAS_R_64F && Ctype == SSBLAS_R_64F) return TT_F64;
    if (Atype == SSBLAS_R_32F && Btype == SSBLAS_R_32F && Ctype == SSBLAS_R_32F) return TT_F32;
    return TT_NONE;
}

// Arithmetic form for a type triple and computeType; -1 = INVALID_VALUE.
// DEFAULT/32I/32F are ignored, as in the reference
// (ssblasGemmBatchedEx_ALLDATASIZESAME.hpp:46-49); appended values must name
// a form that exists for the triple.
int select_form(TypeTriple tt, ssblasComputeType_t ct, int* ozaki_slices)
{
    int c = (int)ct;
    *ozaki_slices = 0;
    if (c == SSBLAS_COMPUTE_DEFAULT_TYPE || c == SSBLAS_COMPUTE_32I || c == SSBLAS_COMPUTE_32F) {
        return tt == TT_I8 ? FORM_I8 : tt == TT_F32 ? FORM_F32 : FORM_F64;
    }
    if (c == SSBLAS_COMPUTE_64F) return tt == TT_F64 ? FORM_F64 : -1;
    if (c == SSBLAS_COMPUTE_32F_FAST_BF16X3) return tt == TT_F32 ? FORM_BF16X3 : -1;
    if (c == SSBLAS_COMPUTE_32F_EMULATED_BF16X6) return tt == TT_F32 ? FORM_BF16X6 : -1;
    if (c == SSBLAS_COMPUTE_64F_EMULATED_OZAKI) return tt == TT_F64 ? FORM_OZAKI : -1;
    if (c >= SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S3 && c <= SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S9) {
        *ozaki_slices = c & 0xff;
        return tt == TT_F64 ? FORM_OZAKI : -1;
    }
    return -1;
}

static long stored_rows(ssblasOperation_t op, long rows_of_op, long cols_of_op)
{
    return op == SSBLAS_OP_N ? rows_of_op : cols_of_op;
}

ssblasStatus_t check_shape(ssblasOperation_t transa, ssblasOperation_t transb, long m, long n, long k,
                           long lda, long ldb, long ldc, long batchCount)
{
    // fp: reference ALLDATASIZESAME.hpp:50-120; int8 OP_C is undefined in the
    // reference (BatchMat::ptr returns NULL) -- rejected here.
    if ((transa != SSBLAS_OP_N && transa != SSBLAS_OP_T) || (transb != SSBLAS_OP_N && transb != SSBLAS_OP_T))
        return SSBLAS_STATUS_INVALID_VALUE;
    // Not checked by the reference (README.md:155): rejected here.
    if (m < 0 || n < 0 || k < 0 || batchCount < 0) return SSBLAS_STATUS_INVALID_VALUE;
    // Zero tasks: reference Launcher/launcher.hpp:174-176.
    if (m == 0 || n == 0 || batchCount == 0) return SSBLAS_STATUS_INTERNAL_ERROR;
    long ra = stored_rows(transa, m, k), rb = stored_rows(transb, k, n);
    if (lda < (ra > 1 ? ra : 1) || ldb < (rb > 1 ? rb : 1) || ldc < (m > 1 ? m : 1))
        return SSBLAS_STATUS_INVALID_VALUE;
    // Sizes beyond 32-bit tile indexing.
    if (m > INT_MAX || n > INT_MAX || k > INT_MAX || batchCount > INT_MAX) return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;
    return SSBLAS_STATUS_SUCCESS;
}

bool device_accessible(const void* p)
{
    if (p == nullptr) return false;
    hipPointerAttribute_t attr;
    memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeHost || attr.type == hipMemoryTypeUnified;
}

size_t workspace_limit()
{
    // SSBLAS_WORKSPACE_LIMIT=<bytes>[K|M|G]: largest device workspace the
    // library may allocate for one call or plan; beyond it the call returns
    // SSBLAS_STATUS_ALLOC_FAILED before any device work.
    const char* e = getenv("SSBLAS_WORKSPACE_LIMIT");
    if (e == nullptr || *e == 0) return SIZE_MAX;
    char* end = nullptr;
    unsigned long long v = strtoull(e, &end, 10);
    if (end != nullptr) {
        if (*end == 'K' || *end == 'k') v <<= 10;
        else if (*end == 'M' || *end == 'm') v <<= 20;
        else if (*end == 'G' || *end == 'g') v <<= 30;
    }
    return (size_t)v;
}

ssblasStatus_t hip_status(hipError_t e)
{
    if (e == hipSuccess) return SSBLAS_STATUS_SUCCESS;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SSBLAS_STATUS_ALLOC_FAILED;
    return SSBLAS_STATUS_INTERNAL_ERROR;
}

void fill_scalars(TypeTriple tt, const void* alpha, const void* beta, GemmProblem& g)
{
    // Read by value as the reference does (..._ALLDATASIZESAME_impl.hpp:146-147).
    if (tt == TT_F32) {
        float a = *static_cast<const float*>(alpha), b = *static_cast<const float*>(beta);
        g.af = a; g.bf = b; g.ad = a; g.bd = b; g.ai = 0; g.bi = 0;
        g.beta0 = (b == 0.0f);
    } else if (tt == TT_F64) {
        double a = *static_cast<const double*>(alpha), b = *static_cast<const double*>(beta);
        g.af = (float)a; g.bf = (float)b; g.ad = a; g.bd = b; g.ai = 0; g.bi = 0;
        g.beta0 = (b == 0.0);
    } else {
        int32_t a = *static_cast<const int32_t*>(alpha), b = *static_cast<const int32_t*>(beta);
        g.ai = a; g.bi = b; g.af = (float)a; g.bf = (float)b; g.ad = a; g.bd = b;
        g.beta0 = (b == 0);
    }
}

// ---------------------------------------------------------------------------
// The batched call
// ---------------------------------------------------------------------------
ssblasStatus_t prepare_call(ssblasOperation_t transa, ssblasOperation_t transb, long m, long n, long k,
                            const void* alpha, const void* const A[], ssblasDataType_t Atype, long lda,
                            const void* const B[], ssblasDataType_t Btype, long ldb, const void* beta,
                            void* const C[], ssblasDataType_t Ctype, long ldc, long batchCount,
                            ssblasComputeType_t computeType, ssblasGemmAlgo_t algo, size_t ws_cap,
                            PreparedCall* pc)
{
    if (getenv("SSBLAS_GEMMBATCHEDEX_DEBUG") != nullptr)
        debug_dump(transa, transb, m, n, k, Atype, lda, Btype, ldb, Ctype, ldc, batchCount, computeType, algo);
    pc->run = false;
    TypeTriple tt = classify_types(Atype, Btype, Ctype);
    if (tt == TT_NONE) return SSBLAS_STATUS_INVALID_VALUE;
    ssblasStatus_t st = check_shape(transa, transb, m, n, k, lda, ldb, ldc, batchCount);
    if (st != SSBLAS_STATUS_SUCCESS) return st;
    int slices = 0;
    int form = select_form(tt, computeType, &slices);
    if (form < 0) return SSBLAS_STATUS_INVALID_VALUE;
    if (alpha == nullptr || beta == nullptr) return SSBLAS_STATUS_INVALID_VALUE;
    if (k == 0) return SSBLAS_STATUS_SUCCESS;  // no K block: C unchanged (reference ..._BMKN.hpp:200)

    GemmProblem& g = pc->g;
    g = GemmProblem{};
    g.ta = transa == SSBLAS_OP_T;
    g.tb = transb == SSBLAS_OP_T;
    g.m = (int)m;
    g.n = (int)n;
    g.k = (int)k;
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    g.batch = (int)batchCount;
    g.A = A;
    g.B = B;
    g.C = C;
    fill_scalars(tt, alpha, beta, g);
    g.arith = form == FORM_I8 ? K_I8 : form == FORM_F32 ? K_F32 : form == FORM_F64 ? K_F64
            : form == FORM_BF16X3 ? K_BF16X3 : form == FORM_BF16X6 ? K_BF16X6 : K_F64;
    pc->form = form;
    pc->slices = slices;
    // Device workspace (the Ozaki form only) beyond the cap.
    if (form == FORM_OZAKI && ozaki_workspace_bytes(g, slices) > ws_cap) return SSBLAS_STATUS_ALLOC_FAILED;
    if (!device_accessible(A) || !device_accessible(B) || !device_accessible(C))
        return SSBLAS_STATUS_INVALID_VALUE;
    pc->run = true;
    return SSBLAS_STATUS_SUCCESS;
}

ssblasStatus_t run_prepared(const PreparedCall& pc, hipStream_t stream, WsGet ws_get)
{
    hipError_t e;
    if (pc.form == FORM_OZAKI) {
        ssblasStatus_t st = SSBLAS_STATUS_SUCCESS;
        e = launch_ozaki(pc.g, pc.slices, stream, ws_get, &st);
        if (st != SSBLAS_STATUS_SUCCESS) return st;
    } else {
        e = launch_gemm(pc.g, pick_config(pc.g), stream);
    }
    return hip_status(e);
}

ssblasStatus_t gemm_batched(ssblasOperation_t transa, ssblasOperation_t transb, long m, long n, long k,
                            const void* alpha, const void* const A[], ssblasDataType_t Atype, long lda,
                            const void* const B[], ssblasDataType_t Btype, long ldb, const void* beta,
                            void* const C[], ssblasDataType_t Ctype, long ldc, long batchCount,
                            ssblasComputeType_t computeType, ssblasGemmAlgo_t algo, hipStream_t stream,
                            bool sync)
{
    PreparedCall pc;
    ssblasStatus_t st = prepare_call(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype,
                                     ldc, batchCount, computeType, algo, workspace_limit(), &pc);
    if (st != SSBLAS_STATUS_SUCCESS || !pc.run) return st;
    st = run_prepared(pc, stream, nullptr);
    if (st != SSBLAS_STATUS_SUCCESS) return st;
    if (sync) return hip_status(hipStreamSynchronize(stream));
    return SSBLAS_STATUS_SUCCESS;
}

}  // namespace ssb

// ===========================================================================
// extern "C"
// ===========================================================================
extern "C" {

ssblasStatus_t ssblasGemmBatchedEx(ssblasOperation_t transa, ssblasOperation_t transb, SSINT m, SSINT n,
                                   SSINT k, const void* alpha, const void* const A[], ssblasDataType_t Atype,
                                   const SSINT lda, const void* const B[], ssblasDataType_t Btype,
                                   const SSINT ldb, const void* beta, void* const C[], ssblasDataType_t Ctype,
                                   const SSINT ldc, const SSINT batchCount, ssblasComputeType_t computeType,
                                   ssblasGemmAlgo_t algo)
{
    return ssb::gemm_batched(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype, ldc,
                             batchCount, computeType, algo, nullptr, true);
}

ssblasStatus_t ssblasGemmBatchedEx_64(ssblasOperation_t transa, ssblasOperation_t transb, SSLONG m, SSLONG n,
                                      SSLONG k, const void* alpha, const void* const A[], ssblasDataType_t Atype,
                                      const SSLONG lda, const void* const B[], ssblasDataType_t Btype,
                                      const SSLONG ldb, const void* beta, void* const C[],
                                      ssblasDataType_t Ctype, const SSLONG ldc, const SSLONG batchCount,
                                      ssblasComputeType_t computeType, ssblasGemmAlgo_t algo)
{
    return ssb::gemm_batched(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype, ldc,
                             batchCount, computeType, algo, nullptr, true);
}

ssblasStatus_t ssblasGemmBatchedExAsync(ssblasOperation_t transa, ssblasOperation_t transb, SSINT m, SSINT n,
                                        SSINT k, const void* alpha, const void* const A[],
                                        ssblasDataType_t Atype, const SSINT lda, const void* const B[],
                                        ssblasDataType_t Btype, const SSINT ldb, const void* beta,
                                        void* const C[], ssblasDataType_t Ctype, const SSINT ldc,
                                        const SSINT batchCount, ssblasComputeType_t computeType,
                                        ssblasGemmAlgo_t algo, ssblasStream_t stream)
{
    return ssb::gemm_batched(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype, ldc,
                             batchCount, computeType, algo, reinterpret_cast<hipStream_t>(stream), false);
}

ssblasStatus_t ssblasGetKernelConfig(ssblasOperation_t transa, ssblasOperation_t transb, SSINT m, SSINT n, SSINT k,
                                     ssblasDataType_t Atype, ssblasDataType_t Btype, ssblasDataType_t Ctype,
                                     const SSINT lda, const SSINT ldb, const SSINT ldc, const SSINT batchCount,
                                     ssblasComputeType_t computeType, char* name, SSINT name_len)
{
    if (name == nullptr || name_len < 1) return SSBLAS_STATUS_INVALID_VALUE;
    name[0] = 0;
    ssb::TypeTriple tt = ssb::classify_types(Atype, Btype, Ctype);
    if (tt == ssb::TT_NONE) return SSBLAS_STATUS_INVALID_VALUE;
    ssblasStatus_t st = ssb::check_shape(transa, transb, m, n, k, lda, ldb, ldc, batchCount);
    if (st != SSBLAS_STATUS_SUCCESS) return st;
    int slices = 0;
    const int form = ssb::select_form(tt, computeType, &slices);
    if (form < 0) return SSBLAS_STATUS_INVALID_VALUE;
    ssb::GemmProblem g{};
    g.ta = transa == SSBLAS_OP_T;
    g.tb = transb == SSBLAS_OP_T;
    g.m = (int)m;
    g.n = (int)n;
    g.k = (int)k;
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    g.batch = (int)batchCount;
    const char* what;
    char buf[64];
    if (form == ssb::FORM_OZAKI) {
        what = ssb::ozaki_kernel_name(g);
    } else {
        g.arith = form == ssb::FORM_I8 ? ssb::K_I8 : form == ssb::FORM_F32 ? ssb::K_F32
                : form == ssb::FORM_F64 ? ssb::K_F64 : form == ssb::FORM_BF16X3 ? ssb::K_BF16X3 : ssb::K_BF16X6;
        what = ssb::config_name(ssb::pick_config(g));
    }
    snprintf(buf, sizeof(buf), "%s", what);
    snprintf(name, (size_t)name_len, "%s", buf);
    return SSBLAS_STATUS_SUCCESS;
}

int ssblasGetVersion(void) { return 200; }

}  // extern "C"
