// gemm_dma_kc.hpp -- direct-to-LDS tiles for K-contiguous fp32 / fp64
// operands (op(A) = A^T, op(B) = B), shared by the native fp32, native fp64
// and bf16-split forms.
//
// Each k-block of A and B moves global -> LDS with global_load_lds_dwordx4
// (no staging registers), S stages deep, and the k-blocks of all tiles of a
// workgroup form one stream: the next tile's first blocks are in flight
// while the current tile's epilogue runs.  LDS keeps the input elements as
// they are (4 or 8 bytes): the bf16 forms split each fragment in registers
// after reading it (two values per v_cvt_pk_bf16_f32), so their LDS tile is
// no larger than the native fp32 one and three stages of a 256x128 tile fit.
//
// LDS layout per operand and stage: row = M/N index, 16-byte chunks along k,
// chunk (row, c) at slot row*CP + (c ^ swz(row)) (the swizzle of the int8
// tiles: 8 consecutive rows of one chunk column hit 8 different 16-byte
// bank groups).
//
// MFMA fragments: lane l supplies row l%16 of the tile fragment and, within
// one MFMA k-step of KSTEP elements, the KSTEP/4 consecutive k starting at
// (l/16)*KSTEP/4, i.e. CPL = KSTEP*E/64 whole chunks: chunk (ks*4 + l/16)*CPL
// + q, q < CPL.  A and B use the same assignment, so every product
// a(i,k)*b(k,j) meets whatever the MFMA's internal k order is.
//
// A tile that is not whole and aligned (instance pointer off 16 bytes, or
// reaching past m / n) is computed by the register-staged plane tile
// (tile_job) in four sub-tiles after the stream drains, and the stream
// restarts behind it.
#pragma once
#include "gemm_dma.hpp"

namespace ssb {

template <class AR, int MN, int BK>
struct KcOperand {
    static constexpr int E = AR::E;
    static constexpr int EPC = 16 / E;              // elements per chunk
    static constexpr int CP = BK / EPC;             // chunks per row
    static constexpr int CHUNKS = MN * CP;
    static constexpr int BYTES = CHUNKS * 16;
    static constexpr int CPL = AR::KSTEP * E / 64;  // chunks per lane per MFMA k-step
    static_assert(BK % EPC == 0 && BK % AR::KSTEP == 0, "whole chunks and k-steps");
    static_assert(CPL == 1 || (CPL == 2 && AR::LDS_E != E), "one chunk, or two split into planes");

    __device__ __forceinline__ static int slot(int row, int c) { return row * CP + (c ^ dma_swz<CP>(row)); }

    // Element offset of chunk (row, c) of the k-block at (mn0, k0).
    __device__ __forceinline__ static long chunk_off(int row, int c, int mn0, int k0, long ld)
    {
        return (long)(mn0 + row) * ld + k0 + EPC * c;
    }

    // The NPLANE fragment registers of tile row `row` for MFMA k-step ks.
    __device__ __forceinline__ static void frag(const char* lds, int row, int ks, int g, u32x4* out)
    {
        if constexpr (CPL == 1) {
            out[0] = *reinterpret_cast<const u32x4*>(lds + slot(row, ks * 4 + g) * 16);
        } else {
            const u32x4 c0 = *reinterpret_cast<const u32x4*>(lds + slot(row, (ks * 4 + g) * 2) * 16);
            const u32x4 c1 = *reinterpret_cast<const u32x4*>(lds + slot(row, (ks * 4 + g) * 2 + 1) * 16);
            u32x4 lo[AR::NPLANE], hi[AR::NPLANE];
            AR::to_planes(c0, lo);
            AR::to_planes(c1, hi);
#pragma unroll
            for (int pl = 0; pl < AR::NPLANE; ++pl) out[pl] = u32x4{lo[pl][0], lo[pl][1], hi[pl][0], hi[pl][1]};
        }
    }
};

template <class AR, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int S>
struct KcTile {
    using OA = KcOperand<AR, BM, BK>;
    using OB = KcOperand<AR, BN, BK>;
    static constexpr int NT = 64 * WAVES_M * WAVES_N;
    static constexpr int NWAVE = WAVES_M * WAVES_N;
    static constexpr int STAGE = OA::BYTES + OB::BYTES;
    static constexpr int LDS_BYTES = S * STAGE;
    static constexpr int WTM = BM / WAVES_M, WTN = BN / WAVES_N;
    static constexpr int FM = WTM / 16, FN = WTN / 16;
    static constexpr int INSTR = (OA::CHUNKS + OB::CHUNKS) / 64;  // wave DMA instructions per stage
    static constexpr int PER_WAVE = INSTR / NWAVE;
    static_assert(OA::CHUNKS % 64 == 0 && OB::CHUNKS % 64 == 0, "whole 1 KiB DMA blocks");
    static_assert(INSTR % NWAVE == 0, "DMA instructions split evenly over the waves");
    static_assert(S >= 2, "at least double buffered");
    static_assert(WTM % 16 == 0 && WTN % 16 == 0, "whole 16x16 fragments");

    __device__ __forceinline__ static bool eligible(const TileJob& j)
    {
        constexpr int E = AR::E;
        return j.i0 + BM <= j.m && j.j0 + BN <= j.n && (j.k % BK) == 0 && ((j.lda * E) % 16) == 0 &&
               ((j.ldb * E) % 16) == 0 && aligned16(j.A) && aligned16(j.B);
    }

    __device__ __forceinline__ static void issue_dma(const TileJob& j, int kt, char* base, int wave, int lane_in)
    {
        constexpr int E = AR::E;
        const int k0 = kt * BK;
        const int lane = opaque(lane_in);
#pragma unroll
        for (int q = 0; q < PER_WAVE; ++q) {
            const int inst = wave + q * NWAVE;
            const int slot = inst * 64 + lane;
            const char* src;
            if (slot < OA::CHUNKS) {
                const int row = slot / OA::CP, c = (slot % OA::CP) ^ dma_swz<OA::CP>(row);
                src = j.A + OA::chunk_off(row, c, j.i0, k0, j.lda) * E;
            } else {
                const int s2 = slot - OA::CHUNKS;
                const int row = s2 / OB::CP, c = (s2 % OB::CP) ^ dma_swz<OB::CP>(row);
                src = j.B + OB::chunk_off(row, c, j.j0, k0, j.ldb) * E;
            }
            __builtin_amdgcn_global_load_lds(src, (lds_void_t*)(base + inst * 1024), 16, 0, 0);
        }
    }

    __device__ __forceinline__ static void compute(typename AR::Acc (&acc)[FM][FN], const char* la, int wm, int wn,
                                                   int lane)
    {
        const char* lb = la + OA::BYTES;
        const int g = lane >> 4;
#pragma unroll
        for (int ks = 0; ks < BK / AR::KSTEP; ++ks) {
            u32x4 af[FM][AR::NPLANE], bfg[FN][AR::NPLANE];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) OA::frag(la, wm * WTM + fm * 16 + frag_row<AR>(lane & 15), ks, g, af[fm]);
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) OB::frag(lb, wn * WTN + fn * 16 + (lane & 15), ks, g, bfg[fn]);
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) AR::mma(acc[fm][fn], af[fm], bfg[fn]);
        }
    }

    __device__ __forceinline__ static void epilogue(const TileJob& j, typename AR::Acc (&acc)[FM][FN], int wm, int wn,
                                                    int lane)
    {
        using Out = typename AR::Out;
        Out* C = static_cast<Out*>(j.C);
        const bool c_fast = j.i0 + BM <= j.m && j.j0 + BN <= j.n && ((j.ldc * (long)sizeof(Out)) % 16) == 0 &&
                            aligned16(C);
        const int rb = j.i0 + wm * WTM + 4 * (lane >> 4), cb = j.j0 + wn * WTN + (lane & 15);
        store_tile<AR>(C, j.ldc, acc, j.s, j.beta0 != 0, c_fast, j.m, j.n, [&](int fm, int fn, int& r0, int& c) {
            r0 = rb + fm * 16;
            c = cb + fn * 16;
        });
    }

    __device__ __forceinline__ static void wait_stage(int ahead)
    {
        if (ahead >= S - 2) wait_vm<(S - 2) * PER_WAVE>();
        else wait_vm<0>();
        lds_barrier();
    }
};

// Streaming kernel over (instance, tile) work items of K-contiguous
// operands.  Tiles that cannot take the DMA path are computed as four
// register-staged sub-tiles of (BM/2) x (BN/2) on the same wave grid, with
// k-depth FBK (small sub-tiles keep that rarely taken path from raising the
// kernel's register count).  WPE: waves per SIMD the registers must allow.
template <class AR, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int S, int FBK, int WPE>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N) __attribute__((amdgpu_waves_per_eu(WPE)))
dma_kc_gemm(const KParams p)
{
    constexpr int FBM = BM / 2, FBN = BN / 2, FWM = WAVES_M, FWN = WAVES_N;
    using T = KcTile<AR, BM, BN, BK, WAVES_M, WAVES_N, S>;
    using FS = TileShape<AR, true, true, FBM, FBN, FBK, FWM, FWN>;
    constexpr int SMEM = T::LDS_BYTES > FS::LDS_BYTES ? T::LDS_BYTES : FS::LDS_BYTES;
    __shared__ __attribute__((aligned(1024))) char smem[SMEM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;
    const int tiles_per_inst = p.tiles_m * p.tiles_n;
    const int nk = p.k / BK;  // whole k-blocks (a K tail makes every tile ineligible)
    const long G = gridDim.x;
    auto job_of = [&](long t) {
        const long w = xcd_remap(t, p.total);
        const int inst = (int)(w / tiles_per_inst);
        const int rem = (int)(w % tiles_per_inst);
        return make_job(p, inst, (rem % p.tiles_m) * BM, (rem / p.tiles_m) * BN);
    };

    long t_c = blockIdx.x;
    if (t_c >= p.total) return;
    TileJob cur = job_of(t_c);
    long t_i = t_c;  // issue cursor: tile, k-block, stream position
    int k_i = 0;
    TileJob ij = cur;
    bool i_ok = T::eligible(ij);
    long u_i = 0, u_c = 0;
    auto issue_next = [&]() {
        if (t_i >= p.total || !i_ok) return false;
        T::issue_dma(ij, k_i, smem + (u_i % S) * T::STAGE, wave, lane);
        ++u_i;
        if (++k_i == nk) {
            k_i = 0;
            t_i += G;
            if (t_i < p.total) {
                ij = job_of(t_i);
                i_ok = T::eligible(ij);
            }
        }
        return true;
    };

    typename AR::Acc acc[T::FM][T::FN];
    while (t_c < p.total) {
        if (!T::eligible(cur)) {
            // nothing of this tile was issued (the stream stops in front of it)
            wait_vm<0>();
            for (int si = 0; si < BM; si += FBM)
                for (int sj = 0; sj < BN; sj += FBN) {
                    TileJob s = cur;
                    s.i0 += si;
                    s.j0 += sj;
                    if (s.i0 < s.m && s.j0 < s.n) tile_job<AR, 1, 1, FBM, FBN, FBK, FWM, FWN, 1>(s, smem);
                }
            t_c += G;
            if (t_c >= p.total) break;
            cur = job_of(t_c);
            t_i = t_c;
            k_i = 0;
            ij = cur;
            i_ok = T::eligible(ij);
            u_c = u_i = 0;
            __syncthreads();  // the fallback's last LDS reads are done before the DMA refills
            continue;
        }
        while (u_i < u_c + S - 1 && issue_next()) {
        }
#pragma unroll
        for (int fm = 0; fm < T::FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < T::FN; ++fn) acc[fm][fn] = typename AR::Acc{};
        for (int kt = 0; kt < nk; ++kt) {
            T::wait_stage((int)min((long)(S - 2), u_i - u_c - 1));
            while (u_i < u_c + S && issue_next()) {
            }
            T::compute(acc, smem + (u_c % S) * T::STAGE, wm, wn, lane);
            ++u_c;
        }
        T::epilogue(cur, acc, wm, wn, lane);
        t_c += G;
        if (t_c < p.total) cur = job_of(t_c);
    }
}

}  // namespace ssb
