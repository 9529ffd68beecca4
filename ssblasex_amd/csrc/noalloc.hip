// noalloc.hip -- ssblasGemmBatchedExNoAlloc: the asynchronous batched call
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The only form whose kernels need device memory beyond A, B and C is the
// general Ozaki path (its int8 slices and exponent vectors).  Here that
// workspace is `scratch`, a file-scope device array of SCRATCH_BYTES that
// every GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): at a device's first call its address is looked up
// once with hipGetSymbolAddress while that device is current and kept per
// device, and that address is what the kernels receive.
//
// Every call on one GPU shares that GPU's scratch: calls that overlap on a
// GPU (two streams, two threads) race on it; calls in one stream are
// ordered.  Its contents persist between calls; the Ozaki kernels write
// every byte they later read (slice padding included), so nothing needs
// clearing.
#include <hip/hip_runtime.h>

#include <mutex>

#include "api_internal.hpp"

namespace ssb {

namespace {

constexpr size_t SCRATCH_BYTES = SSBLAS_NOALLOC_SCRATCH_BYTES;  // 64 MiB per device
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of this device's scratch (looked up once per device, with
// the device current).
hipError_t scratch_address(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (ptr == nullptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// Workspace provider over the scratch array (never allocates).
void* scratch_ws(size_t bytes, hipError_t* err)
{
    if (bytes > SCRATCH_BYTES) {
        *err = hipErrorOutOfMemory;
        return nullptr;
    }
    void* p = nullptr;
    *err = scratch_address(&p);
    return *err == hipSuccess ? p : nullptr;
}

}  // namespace

}  // namespace ssb

extern "C" ssblasStatus_t ssblasGemmBatchedExNoAlloc(ssblasOperation_t transa, ssblasOperation_t transb, SSINT m,
                                                      SSINT n, SSINT k, const void* alpha, const void* const A[],
                                                      ssblasDataType_t Atype, const SSINT lda,
                                                      const void* const B[], ssblasDataType_t Btype,
                                                      const SSINT ldb, const void* beta, void* const C[],
                                                      ssblasDataType_t Ctype, const SSINT ldc,
                                                      const SSINT batchCount, ssblasComputeType_t computeType,
                                                      ssblasGemmAlgo_t algo, ssblasStream_t stream)
{
    ssb::PreparedCall pc;
    // a call whose workspace exceeds the scratch array is refused here,
    // before any device work
    ssblasStatus_t st = ssb::prepare_call(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C,
                                          Ctype, ldc, batchCount, computeType, algo, ssb::SCRATCH_BYTES, &pc);
    if (st != SSBLAS_STATUS_SUCCESS || !pc.run) return st;
    return ssb::run_prepared(pc, reinterpret_cast<hipStream_t>(stream), ssb::scratch_ws);
}
