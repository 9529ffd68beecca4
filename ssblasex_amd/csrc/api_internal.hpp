// api_internal.hpp -- host-side helpers shared by the API translation units.
#pragma once
#include <hip/hip_runtime_api.h>

#include "../../include/ssblasBatch.h"
#include "launch.hpp"

namespace ssb {

enum TypeTriple { TT_NONE = -1, TT_I8 = 0, TT_F32 = 1, TT_F64 = 2 };
enum Form { FORM_I8 = 0, FORM_F32, FORM_F64, FORM_BF16X3, FORM_BF16X6, FORM_OZAKI };

TypeTriple classify_types(ssblasDataType_t Atype, ssblasDataType_t Btype, ssblasDataType_t Ctype);
int select_form(TypeTriple tt, ssblasComputeType_t ct, int* ozaki_slices);
ssblasStatus_t check_shape(ssblasOperation_t transa, ssblasOperation_t transb, long m, long n, long k,
                           long lda, long ldb, long ldc, long batchCount);
bool device_accessible(const void* p);
ssblasStatus_t hip_status(hipError_t e);
// Cap on library-owned device workspace per call or plan
// (SSBLAS_WORKSPACE_LIMIT, default unlimited).
size_t workspace_limit();
void fill_scalars(TypeTriple tt, const void* alpha, const void* beta, GemmProblem& g);

// Tile config for a problem: the autotune table when it has an entry for
// the problem's size bucket, else the built-in heuristic.
int pick_config(const GemmProblem& g);
// Autotuned config for the problem's size bucket, -1 when none.
int tuned_config(const GemmProblem& g);

// Workspace of the Ozaki form: the per-instance report (error bound, slice
// count) at the front -- all the fused form (m, n, k <= 32) needs -- then,
// for the general path, the int32 exponents of op(A)'s rows and op(B)'s
// columns, their statistics (OZAKI_ROW_BYTES each), and the int8 slices of
// op(A) and op(B), k-block major: per instance, per k-block of 32, `s`
// planes of rows_pad x 32 bytes (rows and k padded with zero digits).
constexpr int OZAKI_ROW_BYTES = 48;
struct OzakiLayout {
    long rows_pad_a, rows_pad_b, nkb;  // m rounded up to the tile rows, n to 64; k-blocks of 32 (k padded to 64)
    size_t inst_a, inst_b;             // slice bytes per instance
    size_t off_bound, off_s;           // float bound[batch], int8 slices[batch]
    size_t off_e, off_r;               // exponents (A, then B), row statistics (A, then B)
    size_t off_a, off_b, total;        // A's and B's slices; total bytes
};
// Where an Ozaki call left its per-instance report (device pointers).
struct OzakiReportLoc {
    const float* bound = nullptr;
    const int8_t* slices = nullptr;
    long batch = 0;
};
// A device workspace provider: returns a buffer of at least the requested
// bytes (or nullptr with *err set).
using WsGet = void* (*)(size_t bytes, hipError_t* err);

// One batched call after the host-side checks.
struct PreparedCall {
    GemmProblem g;
    int form;    // Form
    int slices;  // Ozaki slice count named by the compute type (0: automatic)
    bool run;    // false: nothing to compute (k == 0)
};
// The status table of the batched entry points, decided on the host (the
// reference's order of checks); a call whose Ozaki workspace would exceed
// ws_cap gets SSBLAS_STATUS_ALLOC_FAILED.
ssblasStatus_t prepare_call(ssblasOperation_t transa, ssblasOperation_t transb, long m, long n, long k,
                            const void* alpha, const void* const A[], ssblasDataType_t Atype, long lda,
                            const void* const B[], ssblasDataType_t Btype, long ldb, const void* beta,
                            void* const C[], ssblasDataType_t Ctype, long ldc, long batchCount,
                            ssblasComputeType_t computeType, ssblasGemmAlgo_t algo, size_t ws_cap,
                            PreparedCall* pc);
// Enqueue a prepared call on `stream`; ws_get == nullptr: the library-owned
// workspace of (device, stream).
ssblasStatus_t run_prepared(const PreparedCall& pc, hipStream_t stream, WsGet ws_get);

// Ozaki-scheme fp64 (slices == 0: pick automatically).  Sets *st to a
// non-SUCCESS status for failures that are not HIP errors.
hipError_t launch_ozaki(const GemmProblem& g, int slices, hipStream_t stream, WsGet ws_get, ssblasStatus_t* st);
// Device side of the Ozaki form (ozaki.hip); ws_get returns a device buffer
// of at least the requested size owned by the caller's stream.
// slices == 0: the automatic slice count.  *rep receives where the
// per-instance report lands.
hipError_t ozaki_run(const GemmProblem& g, int slices, WsGet ws_get, hipStream_t st, OzakiReportLoc* rep);
OzakiLayout ozaki_layout(const GemmProblem& g, int s);
// Device workspace an Ozaki call of this shape needs with s slices (0 = the
// automatic choice), in bytes.
size_t ozaki_workspace_bytes(const GemmProblem& g, int s);
// The fused one-wave Ozaki form applies (m, n, k <= 32).
bool ozaki_fused_shape(const GemmProblem& g);
// Kernel family an Ozaki call of this shape runs ("ozaki_small" for the
// fused one-wave form, "ozaki_mma" otherwise).
const char* ozaki_kernel_name(const GemmProblem& g);
// Rows of op(A) per tile of the general Ozaki product kernel (op(A)'s slices
// are padded to a multiple of it).
inline int ozaki_tile_rows(long m) { return m > 64 ? 128 : 64; }
// Most slices the automatic choice may take.
constexpr int OZAKI_MAX_SLICES = 9;

}  // namespace ssb
