/*
 * ssblas_oracle.c -- CPU oracle for ssblasGemmBatchedEx.  TEST
 * INFRASTRUCTURE ONLY: the parity tests, the smoke check and the bench's
 * CPU-baseline leg load it (libssblas_oracle.so) as the checker.  Nothing
 * in the shipped library calls it.
 *
 * Three groups of functions:
 *
 * 1. oracle_gemm_batched_ex -- a restatement of the reference's
 *    ssblasGemmBatchedEx on HOST arrays of HOST pointers, with the
 *    reference's status table and its exact floating-point operation order:
 *      - type triple dispatch  (Impl/GemmBatchedEx/ssblasGemmBatchedExSelector.hpp:35-107)
 *      - fp op check N/T only  (ssblasGemmBatchedEx_ALLDATASIZESAME.hpp:50-120)
 *      - alpha/beta read by value as float/double/int32
 *                              (ssblasGemmBatchedEx_ALLDATASIZESAME_impl.hpp:146-147)
 *      - K split in blocks of 512 (launch_kernels/..._tune_kernel_BMKN.hpp:200, :380)
 *      - at the first K block: C = 0 if beta == 0, else C = fl(C*beta)
 *                              (..._BMKN.hpp:219-231; CScale/Xgemm_scale_kernel.hpp:147-189)
 *      - per block: acc = 0; acc = fma(a_k, b_k, acc) for k in order
 *                              (GemmCore/Xgemm_kernel_sve_v4x5_1.hpp:233-338;
 *                               asm/SVE/def_sve_asm_float.h:37-46 fused FMLA)
 *      - per block: C = fma(acc, alpha, C)
 *                              (Xgemm_kernel_sve_v4x5_1.hpp:93-127, :603-604;
 *                               def_sve_asm_float.h:50-58 FMAD zd = za + zd*zm)
 *      - int8: sdot / mad / mul are exact modulo 2^32, so the result is
 *        (alpha*sum_k a*b + [beta != 0]*beta*C) mod 2^32 whatever the
 *        blocking; computed here with int64 sums and uint32 wrap.
 *      - k == 0: no K block runs, C is left unchanged (..._BMKN.hpp:200).
 *    Where the reference is undefined (OP_C with int8, m == 0 division by
 *    zero, negative sizes, too-small leading dimensions) this restates the
 *    MI355X library's documented choice, marked "documented deviation".
 *
 * 2. oracle_gemm_dd -- the high-precision FP64 oracle: every C element
 *    alpha*sum_k a*b + beta*c evaluated in double-double (Dot2 of Ogita,
 *    Rump and Oishi, SIAM J. Sci. Comput. 26(6), 2005, built on Knuth's
 *    TwoSum and an fma-based TwoProduct), returned as an unevaluated
 *    hi + lo pair.  Independent of the reference.  oracle_gemm_dd_batched:
 *    the same per element, over a batch, parallel over instances.
 *
 * 3. oracle_naive_baseline_i8 -- the example's baseline loop for int8
 *    (example/include/naive_gemm.hpp:3-155): OpenMP collapse(3) over
 *    (instance, j, i), int accumulator, C = alpha*tmp + beta*C when
 *    beta != 0 else alpha*tmp.  Used as the timed CPU baseline.
 *    oracle_blas_baseline -- the example's fp baseline
 *    (example/include/naive_gemm.hpp:195-226), a per-instance loop of
 *    cblas_sgemm / cblas_dgemm calls into an OpenBLAS opened at run time.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off, so that
 * the compiler neither fuses nor splits any operation written here).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <dlfcn.h>

#include "../include/ssblasBatch_type.h"

#define ORACLE_BLOCK_K 512 /* reference BLOCK_K, ..._tune_kernel_BMKN.hpp:380 */

typedef long idx_t;

/* Element (r, c) of op(X) for a column-major X with leading dimension ld. */
static inline idx_t op_index(ssblasOperation_t op, idx_t r, idx_t c, idx_t ld)
{
    /* Reference Utils/ssblas_BatchMat.hpp:31-39. */
    return (op == SSBLAS_OP_N) ? (r + c * ld) : (r * ld + c);
}

/* --------------------------------------------------------------------- */
/* 1. Restatement of the reference's ssblasGemmBatchedEx                   */
/* --------------------------------------------------------------------- */

static void gemm_f32_instance(ssblasOperation_t ta, ssblasOperation_t tb,
                              idx_t m, idx_t n, idx_t k, float alpha,
                              const float* A, idx_t lda, const float* B, idx_t ldb,
                              float beta, float* C, idx_t ldc)
{
    for (idx_t k0 = 0; k0 < k; k0 += ORACLE_BLOCK_K) {
        idx_t kb = (k - k0 < ORACLE_BLOCK_K) ? (k - k0) : ORACLE_BLOCK_K;
        for (idx_t j = 0; j < n; ++j) {
            for (idx_t i = 0; i < m; ++i) {
                float* c = &C[i + j * ldc];
                if (k0 == 0) *c = (beta == 0.0f) ? 0.0f : (*c) * beta;
                float acc = 0.0f;
                for (idx_t kk = k0; kk < k0 + kb; ++kk)
                    acc = fmaf(A[op_index(ta, i, kk, lda)], B[op_index(tb, kk, j, ldb)], acc);
                *c = fmaf(acc, alpha, *c);
            }
        }
    }
}

static void gemm_f64_instance(ssblasOperation_t ta, ssblasOperation_t tb,
                              idx_t m, idx_t n, idx_t k, double alpha,
                              const double* A, idx_t lda, const double* B, idx_t ldb,
                              double beta, double* C, idx_t ldc)
{
    for (idx_t k0 = 0; k0 < k; k0 += ORACLE_BLOCK_K) {
        idx_t kb = (k - k0 < ORACLE_BLOCK_K) ? (k - k0) : ORACLE_BLOCK_K;
        for (idx_t j = 0; j < n; ++j) {
            for (idx_t i = 0; i < m; ++i) {
                double* c = &C[i + j * ldc];
                if (k0 == 0) *c = (beta == 0.0) ? 0.0 : (*c) * beta;
                double acc = 0.0;
                for (idx_t kk = k0; kk < k0 + kb; ++kk)
                    acc = fma(A[op_index(ta, i, kk, lda)], B[op_index(tb, kk, j, ldb)], acc);
                *c = fma(acc, alpha, *c);
            }
        }
    }
}

static void gemm_i8_instance(ssblasOperation_t ta, ssblasOperation_t tb,
                             idx_t m, idx_t n, idx_t k, int32_t alpha,
                             const int8_t* A, idx_t lda, const int8_t* B, idx_t ldb,
                             int32_t beta, int32_t* C, idx_t ldc)
{
    if (k <= 0) return; /* no K block: C unchanged */
    for (idx_t j = 0; j < n; ++j) {
        for (idx_t i = 0; i < m; ++i) {
            int64_t s = 0;
            for (idx_t kk = 0; kk < k; ++kk)
                s += (int64_t)A[op_index(ta, i, kk, lda)] * (int64_t)B[op_index(tb, kk, j, ldb)];
            uint32_t r = (uint32_t)alpha * (uint32_t)(uint64_t)s;
            if (beta != 0) r += (uint32_t)beta * (uint32_t)C[i + j * ldc];
            C[i + j * ldc] = (int32_t)r;
        }
    }
}

static idx_t stored_rows(ssblasOperation_t op, idx_t rows_of_op, idx_t cols_of_op)
{
    return (op == SSBLAS_OP_N) ? rows_of_op : cols_of_op;
}

/*
 * Argument checks in the order the MI355X library applies them.  Returns
 * SSBLAS_STATUS_SUCCESS when the call should compute.
 */
ssblasStatus_t oracle_validate(ssblasOperation_t transa, ssblasOperation_t transb,
                               idx_t m, idx_t n, idx_t k,
                               ssblasDataType_t Atype, idx_t lda,
                               ssblasDataType_t Btype, idx_t ldb,
                               ssblasDataType_t Ctype, idx_t ldc, idx_t batchCount)
{
    int f32 = (Atype == SSBLAS_R_32F && Btype == SSBLAS_R_32F && Ctype == SSBLAS_R_32F);
    int f64 = (Atype == SSBLAS_R_64F && Btype == SSBLAS_R_64F && Ctype == SSBLAS_R_64F);
    int i8 = (Atype == SSBLAS_R_8I && Btype == SSBLAS_R_8I && Ctype == SSBLAS_R_32I);
    /* Reference ssblasGemmBatchedExSelector.hpp:105-107. */
    if (!(f32 || f64 || i8)) return SSBLAS_STATUS_INVALID_VALUE;
    /* Reference ssblasGemmBatchedEx_ALLDATASIZESAME.hpp:85-88, :117-120 for
     * fp; int8 OP_C is unchecked there (undefined) -- documented deviation. */
    if ((transa != SSBLAS_OP_N && transa != SSBLAS_OP_T) ||
        (transb != SSBLAS_OP_N && transb != SSBLAS_OP_T))
        return SSBLAS_STATUS_INVALID_VALUE;
    /* Negative sizes: unchecked in the reference (README.md:155) --
     * documented deviation. */
    if (m < 0 || n < 0 || k < 0 || batchCount < 0) return SSBLAS_STATUS_INVALID_VALUE;
    /* Zero tasks: Launcher/launcher.hpp:174-176 (n == 0, batchCount == 0);
     * m == 0 divides by zero in ..._BMKN.hpp:376,382 -- documented choice. */
    if (m == 0 || n == 0 || batchCount == 0) return SSBLAS_STATUS_INTERNAL_ERROR;
    /* Leading dimensions: unchecked in the reference -- documented deviation. */
    idx_t ra = stored_rows(transa, m, k), rb = stored_rows(transb, k, n);
    if (lda < (ra > 1 ? ra : 1) || ldb < (rb > 1 ? rb : 1) || ldc < (m > 1 ? m : 1))
        return SSBLAS_STATUS_INVALID_VALUE;
    return SSBLAS_STATUS_SUCCESS;
}

ssblasStatus_t oracle_gemm_batched_ex(ssblasOperation_t transa, ssblasOperation_t transb,
                                      idx_t m, idx_t n, idx_t k, const void* alpha,
                                      const void* const A[], ssblasDataType_t Atype, idx_t lda,
                                      const void* const B[], ssblasDataType_t Btype, idx_t ldb,
                                      const void* beta, void* const C[], ssblasDataType_t Ctype,
                                      idx_t ldc, idx_t batchCount, int nthreads)
{
    ssblasStatus_t st = oracle_validate(transa, transb, m, n, k, Atype, lda, Btype, ldb,
                                        Ctype, ldc, batchCount);
    if (st != SSBLAS_STATUS_SUCCESS) return st;
    if (nthreads < 1) nthreads = 1;
    if (Atype == SSBLAS_R_32F) {
        float a = *(const float*)alpha, b = *(const float*)beta;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
        for (idx_t i = 0; i < batchCount; ++i)
            gemm_f32_instance(transa, transb, m, n, k, a, (const float*)A[i], lda,
                              (const float*)B[i], ldb, b, (float*)C[i], ldc);
    } else if (Atype == SSBLAS_R_64F) {
        double a = *(const double*)alpha, b = *(const double*)beta;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
        for (idx_t i = 0; i < batchCount; ++i)
            gemm_f64_instance(transa, transb, m, n, k, a, (const double*)A[i], lda,
                              (const double*)B[i], ldb, b, (double*)C[i], ldc);
    } else {
        int32_t a = *(const int32_t*)alpha, b = *(const int32_t*)beta;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
        for (idx_t i = 0; i < batchCount; ++i)
            gemm_i8_instance(transa, transb, m, n, k, a, (const int8_t*)A[i], lda,
                             (const int8_t*)B[i], ldb, b, (int32_t*)C[i], ldc);
    }
    return SSBLAS_STATUS_SUCCESS;
}

/* --------------------------------------------------------------------- */
/* 2. Double-double FP64 oracle                                            */
/* --------------------------------------------------------------------- */

static inline void two_sum(double a, double b, double* s, double* e)
{
    double x = a + b;
    double z = x - a;
    *e = (a - (x - z)) + (b - z);
    *s = x;
}

static inline void two_prod(double a, double b, double* p, double* e)
{
    double x = a * b;
    *e = fma(a, b, -x);
    *p = x;
}

/* One element of the double-double oracle: (*hi, *lo) ~= alpha * sum_k
 * op(A)(i,kk) * op(B)(kk,j) + beta * C0(i,j). */
static void dd_element(ssblasOperation_t ta, ssblasOperation_t tb, idx_t i, idx_t j, idx_t k,
                       double alpha, const void* A, ssblasDataType_t type, idx_t lda,
                       const void* B, idx_t ldb, double beta, const void* C0, idx_t ldc,
                       double* hi, double* lo)
{
    double s = 0.0, t = 0.0; /* running (s + t) */
    for (idx_t kk = 0; kk < k; ++kk) {
        idx_t ia = op_index(ta, i, kk, lda), ib = op_index(tb, kk, j, ldb);
        double a, b;
        if (type == SSBLAS_R_32F) { a = ((const float*)A)[ia]; b = ((const float*)B)[ib]; }
        else if (type == SSBLAS_R_64F) { a = ((const double*)A)[ia]; b = ((const double*)B)[ib]; }
        else { a = ((const int8_t*)A)[ia]; b = ((const int8_t*)B)[ib]; }
        double h, r, q;
        two_prod(a, b, &h, &r);
        two_sum(s, h, &s, &q);
        t += q + r;
    }
    /* (s, t) * alpha */
    double ph, pe;
    two_prod(s, alpha, &ph, &pe);
    pe = fma(t, alpha, pe);
    if (beta != 0.0) {
        double c;
        if (type == SSBLAS_R_32F) c = ((const float*)C0)[i + j * ldc];
        else if (type == SSBLAS_R_64F) c = ((const double*)C0)[i + j * ldc];
        else c = ((const int32_t*)C0)[i + j * ldc];
        double bh, be, q;
        two_prod(beta, c, &bh, &be);
        two_sum(ph, bh, &ph, &q);
        pe += q + be;
    }
    two_sum(ph, pe, hi, lo);
}

/*
 * hi[i] + lo[i] ~= alpha * sum_k op(A)(r,kk) * op(B)(kk,c) + beta * C0(r,c)
 * for one instance, with the element type given by `type` (R_32F / R_64F /
 * R_8I with int32 C).  Outputs are dense column-major m x n (ld = m).
 * beta == 0 drops the C0 term entirely (C0 may be garbage then).
 */
void oracle_gemm_dd(ssblasOperation_t ta, ssblasOperation_t tb, idx_t m, idx_t n, idx_t k,
                    double alpha, const void* A, ssblasDataType_t type, idx_t lda,
                    const void* B, idx_t ldb, double beta, const void* C0, idx_t ldc,
                    double* hi, double* lo)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (idx_t j = 0; j < n; ++j)
        for (idx_t i = 0; i < m; ++i)
            dd_element(ta, tb, i, j, k, alpha, A, type, lda, B, ldb, beta, C0, ldc, &hi[i + j * m],
                       &lo[i + j * m]);
}

/*
 * oracle_gemm_dd over a batch given as host pointer arrays, parallel over
 * instances (for checks of whole benchmark batches of small instances):
 * instance b's result goes to hi/lo + b*m*n, dense column-major m x n.
 * Every element is computed by the same dd_element as oracle_gemm_dd.
 */
void oracle_gemm_dd_batched(ssblasOperation_t ta, ssblasOperation_t tb, idx_t m, idx_t n, idx_t k,
                            double alpha, const void* const A[], ssblasDataType_t type, idx_t lda,
                            const void* const B[], idx_t ldb, double beta, const void* const C0[],
                            idx_t ldc, double* hi, double* lo, idx_t batch, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (idx_t bt = 0; bt < batch; ++bt) {
        double* h = hi + bt * m * n;
        double* l = lo + bt * m * n;
        for (idx_t j = 0; j < n; ++j)
            for (idx_t i = 0; i < m; ++i)
                dd_element(ta, tb, i, j, k, alpha, A[bt], type, lda, B[bt], ldb, beta,
                           beta != 0.0 ? C0[bt] : NULL, ldc, &h[i + j * m], &l[i + j * m]);
    }
}

/* --------------------------------------------------------------------- */
/* 3. The example's int8 baseline loop                                     */
/* --------------------------------------------------------------------- */

/*
 * Restatement of example/include/naive_gemm.hpp:3-155 for int8 -> int32:
 * `#pragma omp parallel for collapse(3)` over (instance, j, i), an int
 * accumulator, and C = alpha*tmp + beta*C (beta != 0) or alpha*tmp.
 * Arithmetic in uint32 so that wrap-around is defined.  Host pointers.
 */
void oracle_naive_baseline_i8(ssblasOperation_t ta, ssblasOperation_t tb, idx_t M, idx_t N,
                              idx_t K, int32_t alpha, const int8_t* const A[], idx_t lda,
                              const int8_t* const B[], idx_t ldb, int32_t beta,
                              int32_t* const C[], idx_t ldc, idx_t batchCount, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for collapse(3) num_threads(nthreads)
    for (idx_t bt = 0; bt < batchCount; ++bt) {
        for (idx_t j = 0; j < N; ++j) {
            for (idx_t i = 0; i < M; ++i) {
                const int8_t* a = A[bt];
                const int8_t* b = B[bt];
                uint32_t tmp = 0;
                for (idx_t kk = 0; kk < K; ++kk)
                    tmp += (uint32_t)(int32_t)a[op_index(ta, i, kk, lda)] *
                           (uint32_t)(int32_t)b[op_index(tb, kk, j, ldb)];
                uint32_t r = (uint32_t)alpha * tmp;
                if (beta != 0) r += (uint32_t)beta * (uint32_t)C[bt][i + j * ldc];
                C[bt][i + j * ldc] = (int32_t)r;
            }
        }
    }
}

/*
 * The example's fp baseline (example/include/naive_gemm.hpp:195-226): a
 * serial loop over the instances, one column-major cblas_sgemm /
 * cblas_dgemm call each (same trans, sizes, leading dimensions, alpha and
 * beta), the BLAS library threading each call.  The image has no cblas.h or
 * system BLAS, so the OpenBLAS that ships with scipy is opened by path at
 * run time: its CBLAS entry points carry the "scipy_" prefix, LP64 ints.
 * Returns 0, or -1 when the library or a symbol is missing.
 */
typedef void (*sgemm_fn)(int, int, int, int, int, int, float, const float*, int, const float*, int,
                         float, float*, int);
typedef void (*dgemm_fn)(int, int, int, int, int, int, double, const double*, int, const double*, int,
                         double, double*, int);
typedef void (*set_threads_fn)(int);

int oracle_blas_baseline(const char* libpath, ssblasDataType_t type, ssblasOperation_t ta,
                         ssblasOperation_t tb, idx_t m, idx_t n, idx_t k, double alpha,
                         const void* const A[], idx_t lda, const void* const B[], idx_t ldb,
                         double beta, void* const C[], idx_t ldc, idx_t batch, int nthreads)
{
    static void* h = NULL;
    if (h == NULL) h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
    if (h == NULL) return -1;
    set_threads_fn set_threads = (set_threads_fn)dlsym(h, "scipy_openblas_set_num_threads");
    if (set_threads != NULL && nthreads > 0) set_threads(nthreads);
    const int layout = 102; /* CblasColMajor */
    const int opa = ta == SSBLAS_OP_T ? 112 : 111, opb = tb == SSBLAS_OP_T ? 112 : 111;
    if (type == SSBLAS_R_32F) {
        sgemm_fn f = (sgemm_fn)dlsym(h, "scipy_cblas_sgemm");
        if (f == NULL) return -1;
        for (idx_t bt = 0; bt < batch; ++bt)
            f(layout, opa, opb, (int)m, (int)n, (int)k, (float)alpha, (const float*)A[bt], (int)lda,
              (const float*)B[bt], (int)ldb, (float)beta, (float*)C[bt], (int)ldc);
        return 0;
    }
    if (type == SSBLAS_R_64F) {
        dgemm_fn f = (dgemm_fn)dlsym(h, "scipy_cblas_dgemm");
        if (f == NULL) return -1;
        for (idx_t bt = 0; bt < batch; ++bt)
            f(layout, opa, opb, (int)m, (int)n, (int)k, alpha, (const double*)A[bt], (int)lda,
              (const double*)B[bt], (int)ldb, beta, (double*)C[bt], (int)ldc);
        return 0;
    }
    return -1;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
