// gemm_dma.hpp -- int8 batched GEMM tile with direct-to-LDS loads.
//
// Each k-block of A and B is moved global -> LDS by global_load_lds_dwordx4
// (64 lanes x 16 bytes -> 1 KiB of LDS per wave instruction, no registers),
// S LDS stages deep: while block kt is multiplied, blocks kt+1 .. kt+S-1
// are in flight.  One barrier per k-block.
//
// LDS holds each operand in its own storage orientation, cut into 16-byte
// chunks of 16 consecutive elements:
//   K-contiguous operand (A^T, B):  row = M/N index, chunks along k
//   M/N-contiguous operand (A, B^T): row = k, chunks along M/N
// Chunk (row, c) sits at chunk slot row*CP + (c ^ swz(row)) (CP chunks per
// row).  The XOR swizzle makes every 8 consecutive rows of one chunk column
// land in 8 different 16-byte bank groups, so the fragment reads below are
// conflict free.
//
// MFMA fragments (v_mfma_i32_16x16x64_i8, lane l: row/column l%16, k bytes
// 16*(l/16) .. +15 of the 64-k step):
//   K-contiguous: one ds_read_b128 of chunk (row, 4*step + l/16).
//   M/N-contiguous: two ds_read_b64_tr_b8.  Within each 16-lane group the
//     instruction gives output lane o the bytes M[2b + (o>>3)][o&7], b < 8,
//     where M[r] is the 8 bytes at lane r's address.  Lane r therefore
//     points at row k0 + (r>>1), bytes 8*(r&1) .. +7 of the fragment's
//     16-element chunk, and output lane o receives elements (k0 + b, o),
//     b = 0..7: eight consecutive k of column o.
//
// Tiles that are not fully inside the matrices, unaligned instances and K
// tails take the guarded path: the same LDS layout written from registers
// (element-wise loads where needed, zero outside the matrix).
#pragma once
#include "gemm_kernels.hpp"

namespace ssb {

typedef __attribute__((address_space(3))) void lds_void_t;
typedef int32_t i32x2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) i32x2v lds_i32x2_t;

// chunk-column swizzle for CP chunks per row
template <int CP>
__device__ __forceinline__ int dma_swz(int row)
{
    if constexpr (CP >= 8) return row & 7;
    else if constexpr (CP == 4) return (row >> 1) & 3;
    else if constexpr (CP == 2) return (row >> 2) & 1;
    else return 0;
}

// One operand's k-block in LDS: KC = K-contiguous storage.
template <bool KC, int MN, int BK>
struct DmaOperand {
    static constexpr int ROWS = KC ? MN : BK;        // LDS rows
    static constexpr int CP = (KC ? BK : MN) / 16;   // 16-byte chunks per row
    static constexpr int CHUNKS = ROWS * CP;
    static constexpr int BYTES = CHUNKS * 16;
    static_assert((KC ? BK : MN) % 16 == 0, "16-byte chunks");

    __device__ __forceinline__ static int slot(int row, int c) { return row * CP + (c ^ dma_swz<CP>(row)); }

    // Element offset (in the instance's storage) of chunk (row, c) of the
    // k-block at (mn0, k0).  KC: element (mn, k) at mn*ld + k; MN: at mn + k*ld.
    __device__ __forceinline__ static long chunk_off(int row, int c, int mn0, int k0, long ld)
    {
        if constexpr (KC) return (long)(mn0 + row) * ld + k0 + 16 * c;
        else return (long)(k0 + row) * ld + mn0 + 16 * c;
    }

    // Fragment for MFMA k-step `ks` (64 k) at tile row/column base `fb`.
    __device__ __forceinline__ static u32x4 frag(const char* lds, int fb, int ks, int lane)
    {
        if constexpr (KC) {
            const int row = fb + (lane & 15);
            const int c = 4 * ks + (lane >> 4);
            return *reinterpret_cast<const u32x4*>(lds + slot(row, c) * 16);
        } else {
            const int r = lane & 15, g = lane >> 4;
            const int c = fb >> 4;  // fb is a multiple of 16
            i32x2v v[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int krow = 64 * ks + 16 * g + 8 * h + (r >> 1);
                const char* p = lds + slot(krow, c) * 16 + 8 * (r & 1);
                v[h] = __builtin_amdgcn_ds_read_tr8_b64_v2i32((lds_i32x2_t*)(p));
            }
            return u32x4{(uint32_t)v[0].x, (uint32_t)v[0].y, (uint32_t)v[1].x, (uint32_t)v[1].y};
        }
    }
};

template <bool AKC, bool BKC, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int S>
struct DmaShape {
    using OA = DmaOperand<AKC, BM, BK>;
    using OB = DmaOperand<BKC, BN, BK>;
    static constexpr int NT = 64 * WAVES_M * WAVES_N;
    static constexpr int NWAVE = WAVES_M * WAVES_N;
    static constexpr int STAGE = OA::BYTES + OB::BYTES;
    static constexpr int LDS_BYTES = S * STAGE;
    static constexpr int WTM = BM / WAVES_M, WTN = BN / WAVES_N;
    static constexpr int FM = WTM / 16, FN = WTN / 16;
    static constexpr int CHUNKS = OA::CHUNKS + OB::CHUNKS;
    static constexpr int INSTR = CHUNKS / 64;           // wave instructions per stage
    static constexpr int PER_WAVE = INSTR / NWAVE;      // DMA instructions per wave per stage
    static_assert(OA::CHUNKS % 64 == 0 && OB::CHUNKS % 64 == 0, "whole 1 KiB DMA blocks");
    static_assert(INSTR % NWAVE == 0, "DMA instructions split evenly over the waves");
    static_assert(BK % 64 == 0, "whole int8 MFMA k-steps");
    static_assert(S >= 2, "at least double buffered");
};

__device__ __forceinline__ void wait_lgkm0() { __builtin_amdgcn_s_waitcnt(0xF | (7 << 4) | (0 << 8) | (3 << 14)); }
// A value the compiler cannot see through: address arithmetic derived from
// it is recomputed where it is used instead of being hoisted out of the
// k-loop and kept live in registers across it.
__device__ __forceinline__ int opaque(int x)
{
    __asm__ volatile("" : "+v"(x));
    return x;
}
// Workgroup barrier that also stops the compiler from moving memory
// accesses (LDS reads, LDS DMA) across it, without the fence's vmcnt wait.
__device__ __forceinline__ void lds_barrier()
{
    __asm__ __volatile__("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ __volatile__("" ::: "memory");
}

// ---------------------------------------------------------------------------
// Building blocks shared by the per-tile and the streaming kernels
// ---------------------------------------------------------------------------
template <bool AKC, bool BKC, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int S>
struct DmaTile {
    using SH = DmaShape<AKC, BKC, BM, BN, BK, WAVES_M, WAVES_N, S>;
    using OA = typename SH::OA;
    using OB = typename SH::OB;
    using AR = ArithI8;

    // The whole tile and every k-block lie inside aligned matrices.
    __device__ __forceinline__ static bool eligible(const TileJob& j)
    {
        return j.i0 + BM <= j.m && j.j0 + BN <= j.n && (j.k % BK) == 0 && (j.lda % 16) == 0 && (j.ldb % 16) == 0 &&
               aligned16(j.A) && aligned16(j.B);
    }

    // DMA k-block kt of an eligible tile into stage buffer `base`.
    __device__ __forceinline__ static void issue_dma(const TileJob& j, int kt, char* base, int wave, int lane_in)
    {
        const int k0 = kt * BK;
        const int lane = opaque(lane_in);
#pragma unroll
        for (int q = 0; q < SH::PER_WAVE; ++q) {
            const int inst = wave + q * SH::NWAVE;  // wave instruction index in the stage
            const int slot = inst * 64 + lane;      // chunk slot written by this lane
            const char* src;
            if (slot < OA::CHUNKS) {
                const int row = slot / OA::CP, c = (slot % OA::CP) ^ dma_swz<OA::CP>(row);
                src = j.A + OA::chunk_off(row, c, j.i0, k0, j.lda);
            } else {
                const int s2 = slot - OA::CHUNKS;
                const int row = s2 / OB::CP, c = (s2 % OB::CP) ^ dma_swz<OB::CP>(row);
                src = j.B + OB::chunk_off(row, c, j.j0, k0, j.ldb);
            }
            __builtin_amdgcn_global_load_lds(src, (lds_void_t*)(base + inst * 1024), 16, 0, 0);
        }
    }

    // Guarded register path: same LDS layout, zero outside the matrices.
    __device__ __forceinline__ static void issue_regs(const TileJob& j, int kt, char* base, int tid)
    {
        const int k0 = kt * BK;
        const int kv = j.k - k0;
        const int rows_a = min(BM, j.m - j.i0), rows_b = min(BN, j.n - j.j0);
        for (int slot = tid; slot < SH::CHUNKS; slot += SH::NT) {
            const bool isa = slot < OA::CHUNKS;
            const int s2 = isa ? slot : slot - OA::CHUNKS;
            const int cp = isa ? OA::CP : OB::CP;
            const int row = s2 / cp;
            const int c = (s2 % cp) ^ (isa ? dma_swz<OA::CP>(row) : dma_swz<OB::CP>(row));
            const bool kc = isa ? AKC : BKC;
            const char* X = isa ? j.A : j.B;
            const long ld = isa ? j.lda : j.ldb;
            const int mn0 = isa ? j.i0 : j.j0;
            const int mnv = isa ? rows_a : rows_b;
            // elements e = 0..15 of the chunk: KC -> (mn = row, k = 16c+e); MN -> (k = row, mn = 16c+e)
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            const long off0 = isa ? OA::chunk_off(row, c, mn0, k0, ld) : OB::chunk_off(row, c, mn0, k0, ld);
            bool whole, in_row;
            if (kc) { in_row = row < mnv; whole = in_row && 16 * c + 16 <= kv; }
            else { in_row = row < kv; whole = in_row && 16 * c + 16 <= mnv; }
            if (whole && (reinterpret_cast<uintptr_t>(X + off0) & 15) == 0) {
                u32x4 v = *gbl(reinterpret_cast<const u32x4*>(X + off0));
                w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
            } else if (in_row) {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const bool ok = kc ? (16 * c + e < kv) : (16 * c + e < mnv);
                    const uint32_t b = ok ? (uint32_t)*gbl(reinterpret_cast<const uint8_t*>(X + off0 + e)) : 0u;
                    w[e >> 2] |= b << (8 * (e & 3));
                }
            }
            *reinterpret_cast<u32x4*>(base + slot * 16) = u32x4{w[0], w[1], w[2], w[3]};
        }
    }

    __device__ __forceinline__ static void compute(i32x4 (&acc)[SH::FM][SH::FN], const char* la, int wm, int wn,
                                                   int lane)
    {
        const char* lb = la + OA::BYTES;
#pragma unroll
        for (int ks = 0; ks < BK / 64; ++ks) {
            u32x4 af[SH::FM], bfg[SH::FN];
#pragma unroll
            for (int fm = 0; fm < SH::FM; ++fm) af[fm] = OA::frag(la, wm * SH::WTM + fm * 16, ks, lane);
#pragma unroll
            for (int fn = 0; fn < SH::FN; ++fn) bfg[fn] = OB::frag(lb, wn * SH::WTN + fn * 16, ks, lane);
#pragma unroll
            for (int fm = 0; fm < SH::FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < SH::FN; ++fn) AR::mma(acc[fm][fn], &af[fm], &bfg[fn]);
        }
    }

    // Returns store_tile's count of vector memory instructions (-1: not fixed).
    // last: the workgroup's last tile (its stores are written through L2).
    __device__ __forceinline__ static int epilogue(const TileJob& j, i32x4 (&acc)[SH::FM][SH::FN], int wm, int wn,
                                                   int lane, bool last)
    {
        int32_t* C = static_cast<int32_t*>(j.C);
        const bool c_fast = j.i0 + BM <= j.m && j.j0 + BN <= j.n && ((j.ldc * 4) % 16) == 0 && aligned16(C);
        const int rb = j.i0 + wm * SH::WTM + 4 * (lane >> 4), cb = j.j0 + wn * SH::WTN + (lane & 15);
        return store_tile<AR, SH::FM, SH::FN, true>(C, j.ldc, acc, j.s, j.beta0 != 0, c_fast, j.m, j.n,
                              [&](int fm, int fn, int& r0, int& c) {
                                  r0 = rb + fm * 16;
                                  c = cb + fn * 16;
                              },
                              last);
    }

    __device__ __forceinline__ static void zero(i32x4 (&acc)[SH::FM][SH::FN])
    {
#pragma unroll
        for (int fm = 0; fm < SH::FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < SH::FN; ++fn) acc[fm][fn] = i32x4{0, 0, 0, 0};
    }

    // Wait until at most `ahead` stages beyond the one about to be used are
    // still in flight (DMA), then barrier.
    // Streaming wait for one k-block: at most `ahead` later k-blocks and, when
    // a tile's epilogue was issued after the block, that epilogue's vector
    // memory instructions (ecase 1: F = FM * FN, ecase 2: 2F) may still be in
    // flight; then barrier.  gfx9 retires a wave's loads and stores in
    // order, so without the epilogue term the wait would also wait for the
    // stores of the tile before.  Counts above the 6-bit field wait at 63
    // (waiting longer is always safe).
    template <int A = S - 2>
    __device__ __forceinline__ static void wait_ring(int ahead, int ecase)
    {
        constexpr int F = SH::FM * SH::FN;
        constexpr int N0 = A * SH::PER_WAVE;
        constexpr int N1 = N0 + F < 63 ? N0 + F : 63;
        constexpr int N2 = N0 + 2 * F < 63 ? N0 + 2 * F : 63;
        if constexpr (A == 0) {
            if (ecase == 2) wait_vm<N2>();
            else if (ecase == 1) wait_vm<N1>();
            else wait_vm<0>();
        } else {
            if (ahead == A) {
                if (ecase == 2) wait_vm<N2>();
                else if (ecase == 1) wait_vm<N1>();
                else wait_vm<N0>();
            } else {
                wait_ring<A - 1>(ahead, ecase);
                return;
            }
        }
        lds_barrier();
    }

    __device__ __forceinline__ static void wait_stage(int ahead)
    {
        if (ahead >= S - 2) wait_vm<(S - 2) * SH::PER_WAVE>();
        else wait_vm<0>();
        lds_barrier();
    }
};

// One tile, S-stage pipeline inside the tile (the previous tile's buffers
// must be free: the routine starts with a barrier).
template <bool AKC, bool BKC, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int S>
__device__ __forceinline__ void dma_tile_job(const TileJob& j, char* smem, bool last)
{
    using T = DmaTile<AKC, BKC, BM, BN, BK, WAVES_M, WAVES_N, S>;
    using SH = typename T::SH;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;
    const int nk = (j.k + BK - 1) / BK;
    const bool dma = T::eligible(j);

    i32x4 acc[SH::FM][SH::FN];
    T::zero(acc);
    lds_barrier();  // previous user of smem has finished reading it
#pragma unroll
    for (int s = 0; s < S - 1; ++s) {
        if (s < nk) {
            if (dma) T::issue_dma(j, s, smem + s * SH::STAGE, wave, lane);
            else T::issue_regs(j, s, smem + s * SH::STAGE, tid);
        }
    }
    for (int kt = 0; kt < nk; ++kt) {
        if (dma) {
            T::wait_stage(min(nk - 1 - kt, S - 2));
        } else {
            wait_lgkm0();
            lds_barrier();
        }
        if (kt + S - 1 < nk) {
            if (dma) T::issue_dma(j, kt + S - 1, smem + ((kt + S - 1) % S) * SH::STAGE, wave, lane);
            else T::issue_regs(j, kt + S - 1, smem + ((kt + S - 1) % S) * SH::STAGE, tid);
        }
        T::compute(acc, smem + (kt % S) * SH::STAGE, wm, wn, lane);
    }
    T::epilogue(j, acc, wm, wn, lane, last);
}

template <bool AKC, bool BKC, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int S>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
dma_gemm_i8(const KParams p)
{
    using SH = DmaShape<AKC, BKC, BM, BN, BK, WAVES_M, WAVES_N, S>;
    __shared__ __attribute__((aligned(1024))) char smem[SH::LDS_BYTES];
    const int tiles_per_inst = p.tiles_m * p.tiles_n;
    for (long t = blockIdx.x; t < p.total; t += gridDim.x) {
        const long w = xcd_remap(t, p.total);
        const int inst = (int)(w / tiles_per_inst);
        const int rem = (int)(w % tiles_per_inst);
        const TileJob j = make_job(p, inst, (rem % p.tiles_m) * BM, (rem / p.tiles_m) * BN);
        dma_tile_job<AKC, BKC, BM, BN, BK, WAVES_M, WAVES_N, S>(j, smem, t + gridDim.x >= p.total);
    }
}

// Streaming kernel for problems whose tiles are all whole (m, n, k multiples
// of the tile, leading dimensions multiples of 16): the k-blocks of all the
// workgroup's tiles form one stream, so the next tile's first blocks are in
// flight while the current tile is finished and its C written.  A tile
// whose instance pointers are not 16-byte aligned is run on its own
// (guarded path) after the stream drains, and the stream restarts behind it.
template <bool AKC, bool BKC, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int S>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
dma_stream_gemm_i8(const KParams p)
{
    using T = DmaTile<AKC, BKC, BM, BN, BK, WAVES_M, WAVES_N, S>;
    using SH = typename T::SH;
    __shared__ __attribute__((aligned(1024))) char smem[SH::LDS_BYTES];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;
    const int tiles_per_inst = p.tiles_m * p.tiles_n;
    const int nk = p.k / BK;
    const long G = gridDim.x;
    auto job_of = [&](long t) {
        const long w = xcd_remap(t, p.total);
        const int inst = (int)(w / tiles_per_inst);
        const int rem = (int)(w % tiles_per_inst);
        return make_job(p, inst, (rem % p.tiles_m) * BM, (rem / p.tiles_m) * BN);
    };

    long t_c = blockIdx.x;  // tile being computed
    if (t_c >= p.total) return;
    TileJob cur = job_of(t_c);
    // issue cursor: tile t_i, k-block k_i, stream position u_i
    long t_i = t_c;
    int k_i = 0;
    TileJob ij = cur;
    bool i_ok = T::eligible(ij);
    long u_i = 0, u_c = 0;
    auto issue_next = [&]() {
        if (t_i >= p.total || !i_ok) return false;
        T::issue_dma(ij, k_i, smem + (u_i % S) * SH::STAGE, wave, lane);
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

    constexpr int F = SH::FM * SH::FN;
    long u_e = 0;   // k-blocks u < u_e were issued before the last epilogue's memory instructions
    int ecase = 0;  // that epilogue: 0 drained, 1 F instructions, 2 2F instructions
    i32x4 acc[SH::FM][SH::FN];
    while (t_c < p.total) {
        if (!T::eligible(cur)) {
            // nothing of this tile was issued (the stream stops in front of it)
            wait_vm<0>();
            dma_tile_job<AKC, BKC, BM, BN, BK, WAVES_M, WAVES_N, S>(cur, smem, t_c + G >= p.total);
            t_c += G;
            if (t_c >= p.total) break;
            cur = job_of(t_c);
            t_i = t_c;
            k_i = 0;
            ij = cur;
            i_ok = T::eligible(ij);
            u_c = u_i = 0;
            u_e = 0;
            ecase = 0;
            lds_barrier();
            continue;
        }
        while (u_i < u_c + S - 1 && issue_next()) {
        }
        T::zero(acc);
        for (int kt = 0; kt < nk; ++kt) {
            T::wait_ring((int)min((long)(S - 2), u_i - u_c - 1), u_c < u_e ? ecase : 0);
            while (u_i < u_c + S && issue_next()) {
            }
            T::compute(acc, smem + (u_c % S) * SH::STAGE, wm, wn, lane);
            ++u_c;
        }
        __asm__ __volatile__("" ::: "memory");  // the k-block loads above stay in front of the epilogue
        const int ne = T::epilogue(cur, acc, wm, wn, lane, t_c + G >= p.total);
        __asm__ __volatile__("" ::: "memory");
        if (ne == F) {
            ecase = 1;
        } else if (ne == 2 * F) {
            ecase = 2;
        } else {
            wait_vm<0>();
            ecase = 0;
        }
        u_e = u_i;
        t_c += G;
        if (t_c < p.total) cur = job_of(t_c);
    }
}

}  // namespace ssb
