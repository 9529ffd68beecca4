// launch_impl.hpp -- instantiates the kernels of one arithmetic form for
// every layout and tile config, and sizes the persistent grid from the
// kernel's occupancy.  Included by exactly one .hip file per form.
#pragma once
#include <stdlib.h>
#include "gemm_dma.hpp"
#include "gemm_dma_kc.hpp"
#include "gemm_kernels.hpp"
#include "launch.hpp"
#include "wave_job.hpp"
#include "grouped_dma.hpp"
#include "occupancy.hpp"

namespace ssb {

// k-depth per (form, config) for the LDS-tiled configs: base depth per form
// (int8 64, fp32 16, fp64 16, bf16 32), doubled where the config says so,
// and halved again while the double-buffered tile would not fit the 160 KiB
// of LDS (never below one MFMA k-step).
template <class AR>
constexpr int base_k()
{
    return AR::ID == ARITH_I8 ? 64 : AR::ID == ARITH_F64 ? 16 : AR::ID == ARITH_F32 ? 16 : 32;
}
template <class AR, int BM, int BN, int WM, int WN>
constexpr int fit_k(int bk)
{
    while (bk > AR::KSTEP &&
           2 * AR::NPLANE * (BM + BN) * lds_row_pitch<AR>(bk) > 160 * 1024)
        bk /= 2;
    return bk;
}
template <class AR, int CFG>
struct CfgShape;
#define SSB_CFG(C, BM_, BN_, WM_, WN_, KMUL, ST)                                     \
    template <class AR>                                                               \
    struct CfgShape<AR, C> {                                                          \
        static constexpr int BM = BM_, BN = BN_, WM = WM_, WN = WN_, STAGES = ST;     \
        static constexpr int BK = fit_k<AR, BM_, BN_, WM_, WN_>(base_k<AR>() * KMUL); \
        static constexpr bool FITS = 2 * AR::NPLANE * (BM_ + BN_) * lds_row_pitch<AR>(BK) <= 160 * 1024; \
    };
SSB_CFG(CFG_TILE_64, 64, 64, 2, 2, (AR::ID == ARITH_F32 ? 2 : 1), 1)
SSB_CFG(CFG_TILE_128, 128, 128, 2, 2, 1, 1)
SSB_CFG(CFG_TILE_128x64, 128, 64, 2, 2, (AR::ID == ARITH_F32 ? 2 : 1), 1)
SSB_CFG(CFG_TILE_64x128, 64, 128, 2, 2, (AR::ID == ARITH_F32 ? 2 : 1), 1)
SSB_CFG(CFG_TILE_128_K2, 128, 128, 2, 2, 2, 1)
SSB_CFG(CFG_TILE_128_W8, 128, 128, 4, 2, (AR::ID == ARITH_F32 ? 2 : 1), 1)
SSB_CFG(CFG_TILE_256x128, 256, 128, 4, 2, 1, 1)
SSB_CFG(CFG_TILE_64_S2, 64, 64, 2, 2, (AR::ID == ARITH_F32 ? 2 : 1), 2)
SSB_CFG(CFG_TILE_128_S2, 128, 128, 2, 2, 1, 2)
SSB_CFG(CFG_TILE_128_K2_S2, 128, 128, 2, 2, 2, 2)
SSB_CFG(CFG_TILE_128_W8_S2, 128, 128, 4, 2, (AR::ID == ARITH_F32 ? 2 : 1), 2)
SSB_CFG(CFG_TILE_256x128_S2, 256, 128, 4, 2, 1, 2)
#undef SSB_CFG

// Persistent grid: every tile gets its own workgroup when they all fit at
// once; otherwise as many workgroups as are resident, a multiple of 8 so
// that each workgroup keeps one XCD's share of the work ids.
// SSBLAS_MAX_GRID=<n> caps the grid below that (tests use it to give every
// workgroup several tiles at small sizes).
inline long persistent_grid(long work, int per_cu)
{
    long slots = (long)per_cu * device_cu_count();
    if (const char* e = getenv("SSBLAS_MAX_GRID")) {
        const long cap = atol(e);
        if (cap >= 1 && cap < slots) slots = cap;
    }
    if (work <= slots) return work;
    return slots >= 8 ? (slots / 8) * 8 : slots;
}

template <class AR, bool AKC, bool BKC, int CFG>
hipError_t run_tile(KParams p, hipStream_t st)
{
    using CS = CfgShape<AR, CFG>;
    if constexpr (!CS::FITS) {
        return run_tile<AR, AKC, BKC, CFG_TILE_128>(p, st);  // tile does not fit in LDS for this form
    } else {
    constexpr int NT = 64 * CS::WM * CS::WN;
    auto kern = tile_gemm<AR, AKC, BKC, CS::BM, CS::BN, CS::BK, CS::WM, CS::WN, CS::STAGES>;
    p.tiles_m = (p.m + CS::BM - 1) / CS::BM;
    p.tiles_n = (p.n + CS::BN - 1) / CS::BN;
    p.total = (long)p.batch * p.tiles_m * p.tiles_n;
    int occ = kernel_occupancy(kern, NT);
    long grid = persistent_grid(p.total, occ);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), 0, st, p);
    return hipGetLastError();
    }
}

template <class AR, bool AKC, bool BKC, int BM, int BN, int BK, int WM, int WN, int S>
hipError_t run_dma(KParams p, hipStream_t st)
{
    if constexpr (AR::ID != ARITH_I8) {
        return run_tile<AR, AKC, BKC, CFG_TILE_128>(p, st);  // direct-to-LDS tiles exist for int8 only
    } else {
        constexpr int NT = 64 * WM * WN;
        p.tiles_m = (p.m + BM - 1) / BM;
        p.tiles_n = (p.n + BN - 1) / BN;
        p.total = (long)p.batch * p.tiles_m * p.tiles_n;
        const bool whole = p.m % BM == 0 && p.n % BN == 0 && p.k % BK == 0 && p.lda % 16 == 0 && p.ldb % 16 == 0;
        if (whole && getenv("SSBLAS_NO_STREAM") == nullptr) {
            auto kern = dma_stream_gemm_i8<AKC, BKC, BM, BN, BK, WM, WN, S>;
            int occ = kernel_occupancy(kern, NT);
            long grid = persistent_grid(p.total, occ);
            hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), 0, st, p);
        } else {
            auto kern = dma_gemm_i8<AKC, BKC, BM, BN, BK, WM, WN, S>;
            int occ = kernel_occupancy(kern, NT);
            long grid = persistent_grid(p.total, occ);
            hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), 0, st, p);
        }
        return hipGetLastError();
    }
}

// Direct-to-LDS tiles of K-contiguous fp32 / fp64 operands.  HCFG: the
// register-staged config of the same shape, used for other layouts and for
// problems whose tiles are not whole.  WPE: waves per SIMD to fit.
template <class AR, bool AKC, bool BKC, int BM, int BN, int WM, int WN, int S, int HCFG, int WPE>
hipError_t run_kc(KParams p, hipStream_t st)
{
    if constexpr (AR::ID == ARITH_I8) {
        return run_dma<AR, AKC, BKC, BM, BN, 64, WM, WN, 3>(p, st);  // int8 has its own direct-to-LDS tiles
    } else if constexpr (!(AKC && BKC)) {
        return run_tile<AR, AKC, BKC, HCFG>(p, st);
    } else {
        constexpr int E = AR::E;
        constexpr int BK = 128 / E;  // 128 bytes of k per row and stage
        const bool whole = p.m % BM == 0 && p.n % BN == 0 && p.k % BK == 0 && (p.lda * E) % 16 == 0 &&
                           (p.ldb * E) % 16 == 0;
        if (!whole) return run_tile<AR, AKC, BKC, HCFG>(p, st);
        constexpr int NT = 64 * WM * WN;
        auto kern = dma_kc_gemm<AR, BM, BN, BK, WM, WN, S, base_k<AR>(), WPE>;
        p.tiles_m = p.m / BM;
        p.tiles_n = p.n / BN;
        p.total = (long)p.batch * p.tiles_m * p.tiles_n;
        int occ = kernel_occupancy(kern, NT);
        long grid = persistent_grid(p.total, occ);
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), 0, st, p);
        return hipGetLastError();
    }
}

// One wave per W x W block.  int8, and fp32 at W = 32: buffer loads
// straight into the fragments, k-steps staged two register sets deep, old C
// read before the k-loop (int8 32^3 x 65536: 150 -> 90 us; fp32 32^3 x 65536:
// 163-188 -> 160-166 us).  fp32 at W = 16 (where the staged path held more
// registers and was slower: 64^3 x 128 4.7 -> 5.3 us), the other forms, and
// matrices whose extent does not fit the descriptors' 31-bit offsets:
// guarded element loads.
template <class AR, bool AKC, bool BKC, int W>
hipError_t run_direct(KParams p, hipStream_t st)
{
    p.tiles_m = (p.m + W - 1) / W;
    p.tiles_n = (p.n + W - 1) / W;
    p.total = (long)p.batch * p.tiles_m * p.tiles_n;
    const long blocks = (p.total + 3) / 4;
    if constexpr (AR::ID == ARITH_I8 || (AR::ID == ARITH_F32 && W == 32)) {
        constexpr int E = AR::E;
        const uint32_t abytes = AKC ? wave_extent(p.k, p.m, p.lda, E) : wave_extent(p.m, p.k, p.lda, E);
        const uint32_t bbytes = BKC ? wave_extent(p.k, p.n, p.ldb, E) : wave_extent(p.n, p.k, p.ldb, E);
        if (abytes != 0u && bbytes != 0u) {
            auto kern = wave_gemm<AR, AKC, BKC, W, W, 2>;
            const long grid = persistent_grid(blocks, kernel_occupancy(kern, 256));
            hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, p, abytes, bbytes);
            return hipGetLastError();
        }
    }
    auto kern = direct_gemm<AR, AKC, BKC, W, W>;
    const long grid = persistent_grid(blocks, kernel_occupancy(kern, 256));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, p);
    return hipGetLastError();
}

// BM x TS tiles (NW waves) with k-blocks loaded straight into LDS through
// buffer descriptors (int8, fp32 and its bf16 split forms; any layout,
// leading dimension and alignment).
// Other forms, or matrices whose extent does not fit the descriptors' 31-bit
// offsets: the register-staged tile of the same shape.
template <class AR, bool AKC, bool BKC, int TS, int BM = TS, int NW = 4>
hipError_t run_gdma(KParams p, hipStream_t st)
{
    if constexpr (AR::ID == ARITH_I8 || AR::ID == ARITH_F32 || AR::NPLANE > 1) {
        constexpr int E = AR::E;
        const uint32_t abytes = AKC ? wave_extent(p.k, p.m, p.lda, E) : wave_extent(p.m, p.k, p.lda, E);
        const uint32_t bbytes = BKC ? wave_extent(p.k, p.n, p.ldb, E) : wave_extent(p.n, p.k, p.ldb, E);
        if (abytes != 0u && bbytes != 0u) {
            p.tiles_m = (p.m + BM - 1) / BM;
            p.tiles_n = (p.n + TS - 1) / TS;
            p.total = (long)p.batch * p.tiles_m * p.tiles_n;
            auto kern = gdma_gemm<AR, AKC, BKC, TS, BM, NW>;
            const long grid = persistent_grid(p.total, kernel_occupancy(kern, 64 * NW));
            hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * NW), 0, st, p, abytes, bbytes);
            return hipGetLastError();
        }
    }
    return run_tile<AR, AKC, BKC, BM == 256 ? CFG_TILE_256x128 : TS == 64 ? CFG_TILE_64 : CFG_TILE_128>(p, st);
}

template <class AR, bool AKC, bool BKC>
hipError_t run_cfg(const KParams& p, int cfg, hipStream_t st)
{
    switch (cfg) {
    case CFG_DIRECT_16: return run_direct<AR, AKC, BKC, 16>(p, st);
    case CFG_DIRECT_32: return run_direct<AR, AKC, BKC, 32>(p, st);
    case CFG_TILE_64: return run_tile<AR, AKC, BKC, CFG_TILE_64>(p, st);
    case CFG_TILE_128: return run_tile<AR, AKC, BKC, CFG_TILE_128>(p, st);
    case CFG_TILE_128x64: return run_tile<AR, AKC, BKC, CFG_TILE_128x64>(p, st);
    case CFG_TILE_64x128: return run_tile<AR, AKC, BKC, CFG_TILE_64x128>(p, st);
    case CFG_TILE_128_K2: return run_tile<AR, AKC, BKC, CFG_TILE_128_K2>(p, st);
    case CFG_TILE_128_W8: return run_tile<AR, AKC, BKC, CFG_TILE_128_W8>(p, st);
    case CFG_TILE_256x128: return run_tile<AR, AKC, BKC, CFG_TILE_256x128>(p, st);
    case CFG_TILE_64_S2: return run_tile<AR, AKC, BKC, CFG_TILE_64_S2>(p, st);
    case CFG_TILE_128_S2: return run_tile<AR, AKC, BKC, CFG_TILE_128_S2>(p, st);
    case CFG_TILE_128_K2_S2: return run_tile<AR, AKC, BKC, CFG_TILE_128_K2_S2>(p, st);
    case CFG_TILE_128_W8_S2: return run_tile<AR, AKC, BKC, CFG_TILE_128_W8_S2>(p, st);
    case CFG_TILE_256x128_S2: return run_tile<AR, AKC, BKC, CFG_TILE_256x128_S2>(p, st);
    case CFG_DMA_64: return run_dma<AR, AKC, BKC, 64, 64, 64, 2, 2, 4>(p, st);
    case CFG_DMA_128: return run_dma<AR, AKC, BKC, 128, 128, 64, 2, 2, 4>(p, st);
    case CFG_DMA_128_K2: return run_dma<AR, AKC, BKC, 128, 128, 128, 2, 2, 3>(p, st);
    case CFG_DMA_256x128: return run_dma<AR, AKC, BKC, 256, 128, 64, 4, 2, 3>(p, st);
    case CFG_DMA_128_S3: return run_dma<AR, AKC, BKC, 128, 128, 64, 2, 2, 3>(p, st);
    case CFG_KC_256x128: return run_kc<AR, AKC, BKC, 256, 128, 4, 2, 3, CFG_TILE_256x128, 2>(p, st);
    case CFG_KC_128: return run_kc<AR, AKC, BKC, 128, 128, 2, 2, 2, CFG_TILE_128, 2>(p, st);
    case CFG_DMA_256: return run_dma<AR, AKC, BKC, 256, 256, 64, 2, 4, 4>(p, st);
    case CFG_DMA_256_W16: return run_dma<AR, AKC, BKC, 256, 256, 64, 4, 4, 4>(p, st);
    case CFG_DMA_256_K2: return run_dma<AR, AKC, BKC, 256, 256, 128, 2, 4, 2>(p, st);
    case CFG_GDMA_128: return run_gdma<AR, AKC, BKC, 128>(p, st);
    case CFG_GDMA_64: return run_gdma<AR, AKC, BKC, 64>(p, st);
    case CFG_GDMA_256x128: return run_gdma<AR, AKC, BKC, 128, 256, 8>(p, st);
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
