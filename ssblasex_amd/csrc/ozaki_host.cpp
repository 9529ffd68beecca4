// ozaki_host.cpp -- workspace management for the Ozaki fp64 form.
//
// The slices and exponents live in a device buffer owned by the library, one
// per (device, stream) so that calls on different streams never share it.
// A buffer only grows; growing waits for the device to go idle before the
// old buffer is released.  Calls on one stream are ordered by the stream.
#include <stdlib.h>

#include <map>
#include <mutex>
#include <utility>

#include "api_internal.hpp"

namespace ssb {

namespace {
struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, Buf> g_bufs;
thread_local hipStream_t t_stream = nullptr;

void* get_ws(size_t bytes, hipError_t* err)
{
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    Buf& b = g_bufs[{dev, t_stream}];
    if (b.bytes >= bytes) return b.ptr;
    if (b.ptr != nullptr) {
        *err = hipDeviceSynchronize();
        if (*err != hipSuccess) return nullptr;
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 8;  // headroom for slightly larger calls
    *err = hipMalloc(&b.ptr, want);
    if (*err != hipSuccess) {
        b.ptr = nullptr;
        return nullptr;
    }
    b.bytes = want;
    return b.ptr;
}
}  // namespace

bool ozaki_fused_shape(const GemmProblem& g)
{
    return g.m <= 32 && g.n <= 32 && g.k <= 32 && getenv("SSBLAS_OZAKI_NO_FUSED") == nullptr;
}

OzakiLayout ozaki_layout(const GemmProblem& g, int s)
{
    OzakiLayout L;
    const long m = g.m, n = g.n, k = g.k;
    const long ra = g.ta ? k : m, ca = g.ta ? m : k;  // storage of A
    const long rb = g.tb ? n : k, cb = g.tb ? k : n;  // storage of B
    L.ld_sa = (ra + 15) / 16 * 16;
    L.ld_sb = (rb + 15) / 16 * 16;
    L.cols_sa = ca;
    L.cols_sb = cb;
    const size_t bytes_a = (size_t)g.batch * s * L.ld_sa * ca;
    const size_t bytes_b = (size_t)g.batch * s * L.ld_sb * cb;
    const size_t bytes_e = (size_t)g.batch * (m + n) * sizeof(int32_t);
    L.off_b = (bytes_a + 255) / 256 * 256;
    L.off_e = L.off_b + (bytes_b + 255) / 256 * 256;
    L.total = L.off_e + bytes_e;
    return L;
}

size_t ozaki_workspace_bytes(const GemmProblem& g, int s)
{
    if (s <= 0) s = OZAKI_DEFAULT_SLICES;
    if (ozaki_fused_shape(g)) return 0;
    return ozaki_layout(g, s).total;
}

const char* ozaki_kernel_name(const GemmProblem& g)
{
    return ozaki_fused_shape(g) ? "ozaki_small" : "ozaki_gemm";
}

hipError_t launch_ozaki(const GemmProblem& g, int slices, hipStream_t stream, WsGet ws_get, ssblasStatus_t* st)
{
    *st = SSBLAS_STATUS_SUCCESS;
    const int s = slices > 0 ? slices : OZAKI_DEFAULT_SLICES;
    if (ws_get == nullptr) {
        t_stream = stream;
        ws_get = get_ws;
    }
    hipError_t e = ozaki_run(g, s, ws_get, stream);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) *st = SSBLAS_STATUS_ALLOC_FAILED;
    return e;
}

}  // namespace ssb
