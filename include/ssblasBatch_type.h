/*
 * ssblasBatch_type.h -- vocabulary of the ssblasGemmBatchedEx C ABI.
 *
 * Every name and numeric value that exists in the reference header
 * (include/ssblasBatch_type.h of hirokitokura/ssblasEx, lines 5-99) is kept
 * identical so that callers compiled against the reference header link and
 * behave unchanged.  Values that the reference does not have are APPENDED
 * after the reference's last enumerator; none of the reference's values moves.
 *
 * Memory convention of this implementation (MI355X / gfx950):
 *   - A[], B[], C[] are device-resident arrays of device pointers
 *     (cuBLAS batched convention), one column-major matrix per entry.
 *   - alpha / beta are host pointers, read once by value, exactly as the
 *     reference reads them (Impl/GemmBatchedEx/
 *     ssblasGemmBatchedEx_ALLDATASIZESAME_impl.hpp:146-147).
 */
#ifndef SSBLASBATCH_TYPE
#define SSBLASBATCH_TYPE

#include <stdio.h>

/* Reference: include/ssblasBatch_type.h:5-6 -- both are `long int`. */
typedef long int SSINT;
typedef long int SSLONG;

/* Reference: include/ssblasBatch_type.h:8-11 (used by the example CLI). */
typedef enum {
    SSBLAS_int_32,
    SSBLAS_int_64
} ssblasIntType_t;

/* Reference: include/ssblasBatch_type.h:14-22. */
typedef enum {
    SSBLAS_STATUS_SUCCESS = 0,
    SSBLAS_STATUS_ALLOC_FAILED = 1,
    SSBLAS_STATUS_INVALID_VALUE = 2,
    SSBLAS_STATUS_INTERNAL_ERROR = 3,
    SSBLAS_STATUS_NOTIMPLEMENTED_ERROR = 4,
    SSBLAS_STATUS_NOT_COMPUTED_ERROR = 5
} ssblasStatus_t;

/* Reference: include/ssblasBatch_type.h:50-54. */
typedef enum {
    SSBLAS_OP_N = 0,
    SSBLAS_OP_T = 1,
    SSBLAS_OP_C = 2
} ssblasOperation_t;

/* Reference: include/ssblasBatch_type.h:58-59. */
typedef signed char SSBLAS_SCHAR;
typedef unsigned char SSBLAS_UCHAR;

/* Reference: include/ssblasBatch_type.h:61-68. */
typedef enum {
    SSBLAS_R_8I = 0,   /* signed char  */
    SSBLAS_R_8U = 1,   /* unsigned char (declared, not a supported type) */
    SSBLAS_R_16F = 2,  /* declared, not a supported type */
    SSBLAS_R_32I = 3,
    SSBLAS_R_32F = 4,
    SSBLAS_R_64F = 5
} ssblasDataType_t;

/*
 * Reference: include/ssblasBatch_type.h:70-74 (values 0..2).  The reference
 * accepts and ignores computeType, so DEFAULT/32I/32F keep selecting the
 * plain native path for every type triple.  Appended values select the
 * split-precision forms; each is only valid for the type triple it names
 * (otherwise SSBLAS_STATUS_INVALID_VALUE).
 */
typedef enum {
    SSBLAS_COMPUTE_DEFAULT_TYPE = 0,
    SSBLAS_COMPUTE_32I = 1,
    SSBLAS_COMPUTE_32F = 2,
    /* ---- appended (no counterpart in the reference) ---- */
    /* fp64 operands, native fp64 MFMA (same as DEFAULT for R_64F).          */
    SSBLAS_COMPUTE_64F = 3,
    /* fp32 operands split into bf16 hi+lo, three bf16 MFMA products
     * (hi*hi + hi*lo + lo*hi) accumulated in fp32.  ~16-bit operands.        */
    SSBLAS_COMPUTE_32F_FAST_BF16X3 = 4,
    /* fp32 operands split into bf16 hi+mid+lo, six bf16 MFMA products
     * (all terms of combined order <= 2) accumulated in fp32. ~24-bit.       */
    SSBLAS_COMPUTE_32F_EMULATED_BF16X6 = 5,
    /* fp64 operands, Ozaki scheme on int8 MFMA: each row of op(A) and column
     * of op(B) is scaled by a power of two and cut into signed 7-bit int8
     * slices; the s(s+1)/2 slice products of the triangle are exact int32
     * GEMMs recombined in fp64.  The slice count is picked automatically.    */
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI = 6,
    /* Same with an explicit slice count s (3..9).                           */
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S3 = 0x103,
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S4 = 0x104,
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S5 = 0x105,
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S6 = 0x106,
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S7 = 0x107,
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S8 = 0x108,
    SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S9 = 0x109
} ssblasComputeType_t;

/* Reference: include/ssblasBatch_type.h:77-79.  Accepted and ignored. */
typedef enum {
    SSBLAS_COMPUTE_DEFAULT = 0
} ssblasGemmAlgo_t;

/* Reference: include/ssblasBatch_type.h:84-89. */
typedef struct {
    char x;
    char y;
    char z;
    char w;
} ssblaschar4;

/* Reference: include/ssblasBatch_type.h:96-99 (internal vocabulary). */
typedef enum {
    SSBLAS_STATUS_COMPUTED = 0,
    SSBLAS_STATUS_NOT_COMPUTED
} ssblasinternalComputeStatus_t;

/*
 * Reference: include/ssblasBatch_type.h:24-47.  Prints the status name on
 * stdout, "SSBLAS_UNKNOWN" for a value outside the table.
 */
static inline void ssblasShowError(ssblasStatus_t retval)
{
    const char* name = "SSBLAS_UNKNOWN";
    switch (retval) {
    case SSBLAS_STATUS_SUCCESS:              name = "SSBLAS_STATUS_SUCCESS"; break;
    case SSBLAS_STATUS_ALLOC_FAILED:         name = "SSBLAS_STATUS_ALLOC_FAILED"; break;
    case SSBLAS_STATUS_INVALID_VALUE:        name = "SSBLAS_STATUS_INVALID_VALUE"; break;
    case SSBLAS_STATUS_INTERNAL_ERROR:       name = "SSBLAS_STATUS_INTERNAL_ERROR"; break;
    case SSBLAS_STATUS_NOTIMPLEMENTED_ERROR: name = "SSBLAS_STATUS_NOTIMPLEMENTED_ERROR"; break;
    case SSBLAS_STATUS_NOT_COMPUTED_ERROR:   name = "SSBLAS_STATUS_NOT_COMPUTED_ERROR"; break;
    default: break;
    }
    printf("%s\n", name);
}

#endif /* SSBLASBATCH_TYPE */
