// ozaki_host.cpp -- workspace and error report of the Ozaki fp64 form.
//
// The per-instance report (error bound, slice count), the slices and the
// exponents live in a device buffer owned by the library, one per (device,
// stream) so that calls on different streams never share it.  A buffer only
// grows; growing waits for the device to go idle before the old buffer is
// released.  Calls on one stream are ordered by the stream.
//
// Every Ozaki call records where its report landed, per (device, stream);
// ssblasGetOzakiReport reads the last one back.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "api_internal.hpp"

namespace ssb {

namespace {
struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, Buf> g_bufs;
std::map<std::pair<int, hipStream_t>, OzakiReportLoc> g_reports;
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

size_t align256(size_t x) { return (x + 255) / 256 * 256; }
}  // namespace

bool ozaki_fused_shape(const GemmProblem& g)
{
    return g.m <= 32 && g.n <= 32 && g.k <= 32 && getenv("SSBLAS_OZAKI_NO_FUSED") == nullptr;
}

OzakiLayout ozaki_layout(const GemmProblem& g, int s)
{
    OzakiLayout L{};
    const long m = g.m, n = g.n, k = g.k;
    const size_t batch = (size_t)g.batch;
    const long tr = ozaki_tile_rows(m);
    L.rows_pad_a = (m + tr - 1) / tr * tr;
    L.rows_pad_b = (n + 63) / 64 * 64;
    L.nkb = (k + 63) / 64 * 2;  // k-blocks of 32, k padded to a multiple of 64
    L.inst_a = (size_t)L.nkb * s * L.rows_pad_a * 32;
    L.inst_b = (size_t)L.nkb * s * L.rows_pad_b * 32;
    L.off_bound = 0;
    L.off_s = align256(batch * sizeof(float));
    L.off_e = L.off_s + align256(batch);
    if (ozaki_fused_shape(g)) {
        L.off_r = L.off_a = L.off_b = L.total = L.off_e;
        return L;
    }
    L.off_r = L.off_e + align256(batch * (m + n) * sizeof(int32_t));
    L.off_a = L.off_r + align256(batch * (m + n) * OZAKI_ROW_BYTES);
    L.off_b = L.off_a + align256(batch * L.inst_a);
    L.total = L.off_b + batch * L.inst_b;
    return L;
}

size_t ozaki_workspace_bytes(const GemmProblem& g, int s)
{
    return ozaki_layout(g, s > 0 ? s : OZAKI_MAX_SLICES).total;
}

const char* ozaki_kernel_name(const GemmProblem& g)
{
    return ozaki_fused_shape(g) ? "ozaki_small" : "ozaki_mma";
}

hipError_t launch_ozaki(const GemmProblem& g, int slices, hipStream_t stream, WsGet ws_get, ssblasStatus_t* st)
{
    *st = SSBLAS_STATUS_SUCCESS;
    if (ws_get == nullptr) {
        t_stream = stream;
        ws_get = get_ws;
    }
    OzakiReportLoc rep;
    hipError_t e = ozaki_run(g, slices, ws_get, stream, &rep);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) *st = SSBLAS_STATUS_ALLOC_FAILED;
    if (e == hipSuccess) {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            std::lock_guard<std::mutex> lk(g_mu);
            g_reports[{dev, stream}] = rep;
        }
    }
    return e;
}

}  // namespace ssb

extern "C" ssblasStatus_t ssblasGetOzakiReport(ssblasStream_t stream, ssblasOzakiReport_t* report, float* bounds,
                                               signed char* slices, SSINT capacity)
{
    if (report == nullptr || capacity < 0) return SSBLAS_STATUS_INVALID_VALUE;
    memset(report, 0, sizeof(*report));
    const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SSBLAS_STATUS_INTERNAL_ERROR;
    ssb::OzakiReportLoc rep;
    {
        std::lock_guard<std::mutex> lk(ssb::g_mu);
        auto it = ssb::g_reports.find({dev, s});
        if (it == ssb::g_reports.end()) return SSBLAS_STATUS_NOT_COMPUTED_ERROR;
        rep = it->second;
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ssb::hip_status(e);
    std::vector<float> b((size_t)rep.batch);
    std::vector<signed char> sl((size_t)rep.batch);
    e = hipMemcpy(b.data(), rep.bound, b.size() * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sl.data(), rep.slices, sl.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return ssb::hip_status(e);
    report->instances = rep.batch;
    report->target = 2.7755575615628914e-17;  // 2^-55
    report->slices_min = 9;
    report->slices_max = 0;
    for (size_t i = 0; i < b.size(); ++i) {
        const double v = b[i];
        if (!(v <= report->bound_max)) report->bound_max = v;  // NaN / inf propagate
        if (!(v <= report->target)) ++report->over_target;
        report->slices_min = std::min<int>(report->slices_min, sl[i]);
        report->slices_max = std::max<int>(report->slices_max, sl[i]);
    }
    const size_t n = std::min<size_t>((size_t)capacity, b.size());
    if (bounds != nullptr) memcpy(bounds, b.data(), n * sizeof(float));
    if (slices != nullptr) memcpy(slices, sl.data(), n);
    return SSBLAS_STATUS_SUCCESS;
}
