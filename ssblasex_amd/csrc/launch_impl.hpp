// launch_impl.hpp -- instantiates the kernels of one arithmetic form for
// every layout and tile config, and sizes the persistent grid from the
// kernel's occupancy.  Included by exactly one .hip file per form.
#pragma once
#include "gemm_kernels.hpp"
#include "launch.hpp"

namespace ssb {

// k-depth per (form, config) for the LDS-tiled configs.
template <class AR, int CFG>
struct TileK {
    static constexpr int value =
        AR::ID == ARITH_I8 ? 64 :
        AR::ID == ARITH_F64 ? 16 :
        AR::ID == ARITH_F32 ? (CFG == CFG_TILE_64 ? 32 : 16) : 32;
};

template <class K>
struct OccCache {
    static int get(K kernel, int threads)
    {
        static int cached[16] = {0};
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
        if (cached[dev] == 0) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess || nb < 1)
                nb = 1;
            cached[dev] = nb;
        }
        return cached[dev];
    }
};

// Persistent grid: every tile gets its own workgroup when they all fit at
// once; otherwise as many workgroups as are resident, a multiple of 8 so
// that each workgroup keeps one XCD's share of the work ids.
inline long persistent_grid(long work, int per_cu)
{
    long slots = (long)per_cu * device_cu_count();
    if (work <= slots) return work;
    return slots >= 8 ? (slots / 8) * 8 : slots;
}

template <class AR, bool AKC, bool BKC, int BM, int BN, int BK>
hipError_t run_tile(KParams p, hipStream_t st)
{
    auto kern = tile_gemm<AR, AKC, BKC, BM, BN, BK, 2, 2>;
    p.tiles_m = (p.m + BM - 1) / BM;
    p.tiles_n = (p.n + BN - 1) / BN;
    p.total = (long)p.batch * p.tiles_m * p.tiles_n;
    int occ = OccCache<decltype(kern)>::get(kern, 256);
    long grid = persistent_grid(p.total, occ);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, p);
    return hipGetLastError();
}

template <class AR, bool AKC, bool BKC, int W>
hipError_t run_direct(KParams p, hipStream_t st)
{
    auto kern = direct_gemm<AR, AKC, BKC, W, W>;
    p.tiles_m = (p.m + W - 1) / W;
    p.tiles_n = (p.n + W - 1) / W;
    p.total = (long)p.batch * p.tiles_m * p.tiles_n;
    int occ = OccCache<decltype(kern)>::get(kern, 256);
    long blocks = (p.total + 3) / 4;
    long grid = persistent_grid(blocks, occ);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, p);
    return hipGetLastError();
}

template <class AR, bool AKC, bool BKC>
hipError_t run_cfg(const KParams& p, int cfg, hipStream_t st)
{
    switch (cfg) {
    case CFG_DIRECT_16: return run_direct<AR, AKC, BKC, 16>(p, st);
    case CFG_DIRECT_32: return run_direct<AR, AKC, BKC, 32>(p, st);
    case CFG_TILE_64: return run_tile<AR, AKC, BKC, 64, 64, TileK<AR, CFG_TILE_64>::value>(p, st);
    case CFG_TILE_128: return run_tile<AR, AKC, BKC, 128, 128, TileK<AR, CFG_TILE_128>::value>(p, st);
    case CFG_TILE_128x64: return run_tile<AR, AKC, BKC, 128, 64, TileK<AR, CFG_TILE_128x64>::value>(p, st);
    case CFG_TILE_64x128: return run_tile<AR, AKC, BKC, 64, 128, TileK<AR, CFG_TILE_64x128>::value>(p, st);
    default: return hipErrorInvalidValue;
    }
}

template <class AR>
hipError_t launch_form(const GemmProblem& g, int cfg, hipStream_t st)
{
    constexpr int E = AR::E;
    constexpr int EO = sizeof(typename AR::Out);
    KParams p;
    p.A = g.A;
    p.B = g.B;
    p.C = g.C;
    p.lda = g.lda;
    p.ldb = g.ldb;
    p.ldc = g.ldc;
    p.m = g.m;
    p.n = g.n;
    p.k = g.k;
    p.batch = g.batch;
    p.s.af = g.af;
    p.s.bf = g.bf;
    p.s.ad = g.ad;
    p.s.bd = g.bd;
    p.s.ai = g.ai;
    p.s.bi = g.bi;
    p.beta0 = g.beta0 ? 1 : 0;
    p.ld_vec_a = (g.lda * E) % 16 == 0;
    p.ld_vec_b = (g.ldb * E) % 16 == 0;
    p.ld_vec_c = (g.ldc * EO) % 16 == 0;
    p.tiles_m = p.tiles_n = 0;
    p.total = 0;
    // op(A) K-contiguous iff A is transposed; op(B) iff B is not.
    if (!g.ta && !g.tb) return run_cfg<AR, false, true>(p, cfg, st);
    if (!g.ta && g.tb) return run_cfg<AR, false, false>(p, cfg, st);
    if (g.ta && !g.tb) return run_cfg<AR, true, true>(p, cfg, st);
    return run_cfg<AR, true, false>(p, cfg, st);
}

}  // namespace ssb
