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
void fill_scalars(TypeTriple tt, const void* alpha, const void* beta, GemmProblem& g);

// Tile config for a problem: the autotune table when it has an entry for
// the problem's size bucket, else the built-in heuristic.
int pick_config(const GemmProblem& g);
// Autotuned config for the problem's size bucket, -1 when none.
int tuned_config(const GemmProblem& g);

// Ozaki-scheme fp64 (slices == 0: pick automatically).  Sets *st to a
// non-SUCCESS status for failures that are not HIP errors.
hipError_t launch_ozaki(const GemmProblem& g, int slices, hipStream_t stream, ssblasStatus_t* st);
// Device side of the Ozaki form (ozaki.hip); ws_get returns a device buffer
// of at least the requested size owned by the caller's stream.
hipError_t ozaki_run(const GemmProblem& g, int s, void* (*ws_get)(size_t, hipError_t*), hipStream_t st);
// Slice count used when the compute type does not name one.
constexpr int OZAKI_DEFAULT_SLICES = 9;

}  // namespace ssb
