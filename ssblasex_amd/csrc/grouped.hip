// grouped.hip -- the grouped (variable-size, mixed-type) batched GEMM: one
// persistent kernel walks a device work list built on the host.
//
// Work item = one output tile of one instance: its tile shape (32x32,
// 64x64 or 128x128, chosen per instance from the size-bucket table), the
// arithmetic form and the operand layouts of the instance's group.  Items
// are sorted by estimated cost, longest first, and workgroup w takes items
// w, w + G, w + 2G, ... so the long tiles start first and the short ones
// fill in behind them.
#include <hip/hip_runtime.h>

#include "gemm_kernels.hpp"
#include "grouped.hpp"
#include "launch.hpp"

namespace ssb {

template <class AR, int CFG>
struct GTile;
template <class AR> struct GTile<AR, GCFG_32> { static constexpr int BM = 32, BN = 32, BK = (AR::ID == ARITH_I8) ? 64 : (AR::ID == ARITH_F64 ? 16 : 32); };
template <class AR> struct GTile<AR, GCFG_64> { static constexpr int BM = 64, BN = 64, BK = (AR::ID == ARITH_I8) ? 64 : (AR::ID == ARITH_F64 ? 16 : 32); };
template <class AR> struct GTile<AR, GCFG_128> { static constexpr int BM = 128, BN = 128, BK = (AR::ID == ARITH_I8) ? 64 : (AR::ID == ARITH_F64 ? 16 : (AR::ID == ARITH_F32 ? 16 : 32)); };

template <class AR, int CFG>
constexpr int glds()
{
    using T = GTile<AR, CFG>;
    return TileShape<AR, true, true, T::BM, T::BN, T::BK, 2, 2>::LDS_BYTES;
}
template <class AR>
constexpr int glds_max()
{
    return glds<AR, GCFG_128>() > glds<AR, GCFG_64>() ? glds<AR, GCFG_128>() : glds<AR, GCFG_64>();
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int GROUPED_LDS = cmax(cmax(glds_max<ArithI8>(), glds_max<ArithF32>()),
                                 cmax(glds_max<ArithF64>(), cmax(glds_max<ArithBF16X3>(), glds_max<ArithBF16X6>())));

template <class AR, bool AKC, bool BKC>
__device__ __forceinline__ void run_item(int cfg, const TileJob& j, char* smem)
{
    switch (cfg) {
    case GCFG_32: {
        using T = GTile<AR, GCFG_32>;
        tile_job<AR, AKC, BKC, T::BM, T::BN, T::BK, 2, 2>(j, smem);
        break;
    }
    case GCFG_64: {
        using T = GTile<AR, GCFG_64>;
        tile_job<AR, AKC, BKC, T::BM, T::BN, T::BK, 2, 2>(j, smem);
        break;
    }
    default: {
        using T = GTile<AR, GCFG_128>;
        tile_job<AR, AKC, BKC, T::BM, T::BN, T::BK, 2, 2>(j, smem);
        break;
    }
    }
}

template <class AR>
__device__ __forceinline__ void run_form(int layout, int cfg, const TileJob& j, char* smem)
{
    // layout bit 1: op(A) K-contiguous (A transposed); bit 0: op(B) K-contiguous (B not transposed)
    switch (layout) {
    case 0: run_item<AR, false, false>(cfg, j, smem); break;
    case 1: run_item<AR, false, true>(cfg, j, smem); break;
    case 2: run_item<AR, true, false>(cfg, j, smem); break;
    default: run_item<AR, true, true>(cfg, j, smem); break;
    }
}

__global__ void __launch_bounds__(256) grouped_gemm(const GItem* __restrict__ items, int nitems,
                                                     const GInst* __restrict__ insts,
                                                     const GGroup* __restrict__ groups, const void* const* A,
                                                     const void* const* B, void* const* C)
{
    __shared__ __attribute__((aligned(16))) char smem[GROUPED_LDS];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const GItem item = items[it];
        const GInst in = insts[item.inst];
        const GGroup& g = groups[in.group];
        TileJob j;
        j.A = static_cast<const char*>(A[item.inst]);
        j.B = static_cast<const char*>(B[item.inst]);
        j.C = C[item.inst];
        j.lda = in.lda;
        j.ldb = in.ldb;
        j.ldc = in.ldc;
        j.m = in.m;
        j.n = in.n;
        j.k = in.k;
        j.i0 = item.i0;
        j.j0 = item.j0;
        j.s = g.s;
        j.beta0 = g.beta0;
        const int layout = (g.ta ? 2 : 0) | (g.tb ? 0 : 1);
        switch (g.form) {
        case K_I8: run_form<ArithI8>(layout, item.cfg, j, smem); break;
        case K_F32: run_form<ArithF32>(layout, item.cfg, j, smem); break;
        case K_F64: run_form<ArithF64>(layout, item.cfg, j, smem); break;
        case K_BF16X3: run_form<ArithBF16X3>(layout, item.cfg, j, smem); break;
        default: run_form<ArithBF16X6>(layout, item.cfg, j, smem); break;
        }
    }
}

hipError_t launch_grouped(const GroupedDevice& d, hipStream_t st)
{
    static int occ[16] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
    if (occ[dev] == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, grouped_gemm, 256, 0) != hipSuccess || nb < 1) nb = 1;
        occ[dev] = nb;
    }
    long slots = (long)occ[dev] * device_cu_count();
    long grid = d.nitems <= slots ? d.nitems : slots;
    if (grid < 1) return hipSuccess;
    hipLaunchKernelGGL(grouped_gemm, dim3((unsigned)grid), dim3(256), 0, st, d.items, d.nitems, d.insts, d.groups,
                       d.A, d.B, d.C);
    return hipGetLastError();
}

int grouped_lds_bytes() { return GROUPED_LDS; }

}  // namespace ssb
