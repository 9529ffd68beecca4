// grouped_dma.hpp -- the grouped kernel's 64x64 and 128x128 workgroup tiles
// and its one-wave 32x32 items for the int8 and fp32 forms with
// direct-to-LDS buffer loads.
//
// Every k-block of A and B moves global -> LDS with buffer_load_dwordx4 ...
// lds (64 lanes x 16 bytes -> 1 KiB of LDS per wave instruction, no staging
// registers), S stages deep: while block kt is multiplied, blocks kt+1 ..
// kt+S-1 are in flight.  One barrier per k-block.  The loads go through
// buffer descriptors spanning each stored matrix, so they are range-checked
// like the register-staged tiles' loads: a dword past the end of the matrix
// lands in LDS as zero, any leading dimension and alignment is read whole.
//
// LDS keeps each operand in its storage orientation, in 16-byte chunks
// (int8: the layout of the batched direct-to-LDS tiles, DmaOperand; fp32
// K-contiguous: KcOperand; fp32 M/N-contiguous: rows of k, chunks of four
// M/N values, F32MnOperand below).  Each lane of a DMA instruction writes
// the chunk slot of its position, so the global chunk a lane fetches is the
// one the swizzled layout puts at that slot.
//
// What the loads cannot get right by themselves, fixed in LDS by the lane
// that loaded the chunk, after its wave's loads of the block have landed and
// before the barrier that publishes the block (a one-wave item: before the
// wave reads the block):
//  * the k tail of a K-contiguous operand (last k-block): elements at k >= K
//    hold the next row or padding and are cleared (M/N-contiguous chunks at
//    k >= K start past the end of the matrix and are zero already);
//  * int8: the dword holding a matrix's last bytes reaches past its end and
//    reads as zero, so a chunk that crosses the end is read again as the 16
//    bytes ending at the end and shifted (in a k-block whose footprint
//    reaches past the end, and only where a lane of the wave holds such a
//    chunk).
// Rows and columns outside the instance are not masked: they only feed
// outputs that are never stored.
#pragma once
#include "gemm_dma_kc.hpp"
#include "wave_job.hpp"

namespace ssb {

// fp32 M/N-contiguous operand in LDS: row = k, 16-byte chunks of 4
// consecutive M/N values; chunk (row, c) at slot row*CP + (c ^ swz(row)).
// A fragment lane (r = l%16, g = l/16) needs the four k = 16 ks + 4 g + i,
// i < 4, of tile row fb + r (the k assignment of KcOperand's chunks), four
// dword reads from rows that share their swizzle; the swizzle moves the four
// lane groups' rows (k differing in bits 2..3) to different 64-byte bank
// ranges (32 columns, 128-byte rows: the two lane groups of each half-wave).
template <int MN, int BK>
struct F32MnOperand {
    static constexpr int CP = MN / 4;
    static constexpr int CHUNKS = BK * CP;
    static constexpr int BYTES = CHUNKS * 16;
    static_assert((CP == 8 || CP >= 16) && BK % 16 == 0, "two or four 64-byte column groups, whole k-steps");

    __device__ __forceinline__ static int swz(int row)
    {
        return CP >= 16 ? ((row >> 2) & 3) << 2 : ((row >> 2) & 1) << 2;
    }
    __device__ __forceinline__ static int slot(int row, int c) { return row * CP + (c ^ swz(row)); }
    __device__ __forceinline__ static long chunk_off(int row, int c, int mn0, int k0, long ld)
    {
        return (long)(k0 + row) * ld + mn0 + 4 * c;
    }
    __device__ __forceinline__ static u32x4 frag(const char* lds, int fb, int ks, int lane)
    {
        const int mn = fb + (lane & 15), g = lane >> 4;
        const int c = mn >> 2, w = mn & 3;
        uint32_t v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 16 * ks + 4 * g + i;
            v[i] = *reinterpret_cast<const uint32_t*>(lds + slot(row, c) * 16 + 4 * w);
        }
        return u32x4{v[0], v[1], v[2], v[3]};
    }
};

// One operand of a grouped DMA tile: the LDS layout for (form, layout), the
// chunk a slot holds, the k range a K-contiguous chunk covers, and fragment
// reads for one MFMA k-step.
template <class AR, bool KC, int MN, int BK>
struct GDmaOp;

template <bool KC, int MN, int BK>
struct GDmaOp<ArithI8, KC, MN, BK> {
    using L = DmaOperand<KC, MN, BK>;
    static constexpr int CP = L::CP, CHUNKS = L::CHUNKS, BYTES = L::BYTES, EPC = 16;
    __device__ __forceinline__ static void coords(int s, int& row, int& c)
    {
        row = s / CP;
        c = (s % CP) ^ dma_swz<CP>(row);
    }
    __device__ __forceinline__ static long off(int row, int c, int mn0, int k0, long ld)
    {
        return L::chunk_off(row, c, mn0, k0, ld);
    }
    __device__ __forceinline__ static void frag(const char* lds, int fb, int ks, int lane, u32x4* out)
    {
        out[0] = L::frag(lds, fb, ks, lane);
    }
};

template <int MN, int BK>
struct GDmaOp<ArithF32, true, MN, BK> {
    using L = KcOperand<ArithF32, MN, BK>;
    static constexpr int CP = L::CP, CHUNKS = L::CHUNKS, BYTES = L::BYTES, EPC = 4;
    __device__ __forceinline__ static void coords(int s, int& row, int& c)
    {
        row = s / CP;
        c = (s % CP) ^ dma_swz<CP>(row);
    }
    __device__ __forceinline__ static long off(int row, int c, int mn0, int k0, long ld)
    {
        return 4 * L::chunk_off(row, c, mn0, k0, ld);
    }
    __device__ __forceinline__ static void frag(const char* lds, int fb, int ks, int lane, u32x4* out)
    {
        L::frag(lds, fb + (lane & 15), ks, lane >> 4, out);
    }
};

template <int MN, int BK>
struct GDmaOp<ArithF32, false, MN, BK> {
    using L = F32MnOperand<MN, BK>;
    static constexpr int CP = L::CP, CHUNKS = L::CHUNKS, BYTES = L::BYTES, EPC = 4;
    __device__ __forceinline__ static void coords(int s, int& row, int& c)
    {
        row = s / CP;
        c = (s % CP) ^ L::swz(row);
    }
    __device__ __forceinline__ static long off(int row, int c, int mn0, int k0, long ld)
    {
        return 4 * L::chunk_off(row, c, mn0, k0, ld);
    }
    __device__ __forceinline__ static void frag(const char* lds, int fb, int ks, int lane, u32x4* out)
    {
        out[0] = L::frag(lds, fb, ks, lane);
    }
};

// The bf16 split forms keep fp32 in LDS in the fp32 layouts and cut each
// fragment into its bf16 planes after reading it: a lane's eight k of one
// 32-deep MFMA k-step are two 16-byte chunks (K-contiguous, KcOperand's
// two-chunk read) or eight dwords from k-rows 32 ks + 8 g + i (M/N-contiguous).
template <int NP, int ORD, int IDV, int MN, int BK>
struct GDmaOp<ArithBF16Split<NP, ORD, IDV>, true, MN, BK> {
    using AR = ArithBF16Split<NP, ORD, IDV>;
    using L = KcOperand<AR, MN, BK>;
    static constexpr int CP = L::CP, CHUNKS = L::CHUNKS, BYTES = L::BYTES, EPC = 4;
    __device__ __forceinline__ static void coords(int s, int& row, int& c)
    {
        row = s / CP;
        c = (s % CP) ^ dma_swz<CP>(row);
    }
    __device__ __forceinline__ static long off(int row, int c, int mn0, int k0, long ld)
    {
        return 4 * L::chunk_off(row, c, mn0, k0, ld);
    }
    __device__ __forceinline__ static void frag(const char* lds, int fb, int ks, int lane, u32x4* out)
    {
        L::frag(lds, fb + (lane & 15), ks, lane >> 4, out);
    }
};

template <int NP, int ORD, int IDV, int MN, int BK>
struct GDmaOp<ArithBF16Split<NP, ORD, IDV>, false, MN, BK> {
    using AR = ArithBF16Split<NP, ORD, IDV>;
    using L = F32MnOperand<MN, BK>;
    static constexpr int CP = L::CP, CHUNKS = L::CHUNKS, BYTES = L::BYTES, EPC = 4;
    static_assert(BK % 32 == 0, "whole bf16 k-steps");
    __device__ __forceinline__ static void coords(int s, int& row, int& c)
    {
        row = s / CP;
        c = (s % CP) ^ L::swz(row);
    }
    __device__ __forceinline__ static long off(int row, int c, int mn0, int k0, long ld)
    {
        return 4 * L::chunk_off(row, c, mn0, k0, ld);
    }
    __device__ __forceinline__ static void frag(const char* lds, int fb, int ks, int lane, u32x4* out)
    {
        const int mn = fb + (lane & 15), g = lane >> 4;
        const int c = mn >> 2, w = mn & 3;
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = 32 * ks + 8 * g + i;
            v[i] = *reinterpret_cast<const uint32_t*>(lds + L::slot(row, c) * 16 + 4 * w);
        }
        u32x4 lo[NP], hi[NP];
        AR::to_planes(u32x4{v[0], v[1], v[2], v[3]}, lo);
        AR::to_planes(u32x4{v[4], v[5], v[6], v[7]}, hi);
#pragma unroll
        for (int p = 0; p < NP; ++p) out[p] = u32x4{lo[p][0], lo[p][1], hi[p][0], hi[p][1]};
    }
};

// k-depth and stage count of a grouped DMA tile per form and tile size:
// 16 KiB per stage (int8 128x64 / 64x128 k-blocks, fp32 128x16 / 64x32),
// three stages (48 KiB: three workgroups per CU).
template <class AR, int T>
struct GDmaCfg {
    static constexpr bool SPLIT = AR::NPLANE > 1;
    // the bf16 forms' k-step is 32 deep: two stages of 32 KiB (two workgroups per CU)
    static constexpr int BK = SPLIT ? 32 : (AR::ID == ARITH_I8 ? 64 : 16) * (128 / T);
    static constexpr int S = SPLIT ? 2 : 3;
};

template <class AR, bool AKC, bool BKC, int BM, int BN, int BK, int S, int NW = 4>
struct GDmaTile {
    using OA = GDmaOp<AR, AKC, BM, BK>;
    using OB = GDmaOp<AR, BKC, BN, BK>;
    static_assert(NW == 8 || NW == 4 || NW == 1, "4 x 2 or 2 x 2 waves, or one wave over the whole tile");
    static constexpr int NWAVE = NW;
    static constexpr int WAVES_M = NW == 8 ? 4 : NW == 4 ? 2 : 1, WAVES_N = NW == 1 ? 1 : 2;
    static constexpr int WTM = BM / WAVES_M, WTN = BN / WAVES_N;
    static constexpr int FM = WTM / 16, FN = WTN / 16;
    static constexpr int STAGE = OA::BYTES + OB::BYTES;
    static constexpr int LDS_BYTES = S * STAGE;
    static constexpr int A_PW = OA::CHUNKS / 64 / NWAVE;  // A instructions per wave per stage
    static constexpr int B_PW = OB::CHUNKS / 64 / NWAVE;
    static constexpr int PER_WAVE = A_PW + B_PW;
    static_assert(OA::CHUNKS % (64 * NWAVE) == 0 && OB::CHUNKS % (64 * NWAVE) == 0,
                  "whole 1 KiB DMA blocks, split evenly over the waves");
    static_assert(BK % AR::KSTEP == 0 && S >= 2 && S - 2 <= 15, "whole k-steps; stage count");
    static_assert(PER_WAVE * (S - 2) < 64, "vmcnt range");

    // Byte offset (in the stored matrix) of the chunk that LDS slot s of
    // operand O holds for the k-block at k0.
    template <class O>
    __device__ __forceinline__ static uint32_t src(int s, int mn0, int k0, long ld)
    {
        int row, c;
        O::coords(s, row, c);
        return (uint32_t)O::off(row, c, mn0, k0, ld);
    }

    __device__ __forceinline__ static void issue(const TileJob& j, __amdgpu_buffer_rsrc_t ra,
                                                 __amdgpu_buffer_rsrc_t rb, int kt, char* base, int wave,
                                                 int lane_in)
    {
        const int k0 = kt * BK;
        const int lane = opaque(lane_in);
#pragma unroll
        for (int q = 0; q < A_PW; ++q) {
            const int inst = wave + q * NWAVE;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void_t*)(base + inst * 1024), 16,
                                                     (int)src<OA>(inst * 64 + lane, j.i0, k0, j.lda), 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < B_PW; ++q) {
            const int inst = wave + q * NWAVE;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void_t*)(base + OA::BYTES + inst * 1024), 16,
                                                     (int)src<OB>(inst * 64 + lane, j.j0, k0, j.ldb), 0, 0, 0);
        }
    }

    // int8: the chunks this lane loaded for the k-block at k0 that cross the
    // end of the matrix (`ext` bytes), read again as the 16 bytes ending at
    // the end and shifted, and written over the loaded chunk.  Only called
    // for a k-block whose footprint reaches past the end (with a tiny M/N
    // extent and no padding that can be any k-block, not only the last).
    template <class O, int PW>
    __device__ __forceinline__ static void fix_end(char* opbase, __amdgpu_buffer_rsrc_t rs, uint32_t ext, int mn0,
                                                   int k0, long ld, int wave, int lane)
    {
        bool any = false;
#pragma unroll
        for (int q = 0; q < PW; ++q) {
            const uint32_t o = src<O>((wave + q * NWAVE) * 64 + lane, mn0, k0, ld);
            any |= o < ext && o + 16u > ext;
        }
        if (!__any(any)) return;
#pragma unroll
        for (int q = 0; q < PW; ++q) {
            const int s = (wave + q * NWAVE) * 64 + lane;
            const uint32_t o = src<O>(s, mn0, k0, ld);
            if (o < ext && o + 16u > ext) {
                uint32_t d[4] = {0u, 0u, 0u, 0u};
                if (ext < 16u) {
                    for (uint32_t p = o; p < ext; ++p)
                        d[(p - o) >> 2] |= __builtin_amdgcn_raw_buffer_load_b8(rs, (int)p, 0, 0) << (8 * ((p - o) & 3));
                } else {
                    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(ext - 16u), 0, 0);
                    const uint32_t sh = o - (ext - 16u), qq = sh >> 2, r = sh & 3;
                    const uint32_t w[4] = {v[0], v[1], v[2], v[3]};
                    uint32_t t[5];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t a = i + qq;
                        t[i] = a == 0 ? w[0] : a == 1 ? w[1] : a == 2 ? w[2] : a == 3 ? w[3] : 0u;
                    }
                    t[4] = 0u;
#pragma unroll
                    for (int i = 0; i < 4; ++i) d[i] = __builtin_amdgcn_alignbyte(t[i + 1], t[i], r);
                }
                *reinterpret_cast<u32x4*>(opbase + s * 16) = u32x4{d[0], d[1], d[2], d[3]};
            }
        }
    }

    // K-contiguous operand, last k-block (kv < BK valid k from its start):
    // the elements at k >= kv of the chunks this lane loaded cleared.
    template <class O, int PW>
    __device__ __forceinline__ static void clear_k(char* opbase, int kv, int wave, int lane)
    {
#pragma unroll
        for (int q = 0; q < PW; ++q) {
            const int s = (wave + q * NWAVE) * 64 + lane;
            int row, c;
            O::coords(s, row, c);
            const int valid = kv - c * O::EPC;
            if (valid < O::EPC) {
                u32x4* p = reinterpret_cast<u32x4*>(opbase + s * 16);
                const u32x4 v = *p;
                uint32_t d[4] = {v[0], v[1], v[2], v[3]};
                clear_tail<16, AR::E>(d, max(valid, 0));
                *p = u32x4{d[0], d[1], d[2], d[3]};
            }
        }
    }

    __device__ __forceinline__ static void compute(typename AR::Acc (&acc)[FM][FN], const char* la, int wm, int wn,
                                                   int lane)
    {
        const char* lb = la + OA::BYTES;
#pragma unroll
        for (int ks = 0; ks < BK / AR::KSTEP; ++ks) {
            u32x4 af[FM][AR::NPLANE], bfg[FN][AR::NPLANE];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) OA::frag(la, wm * WTM + fm * 16, ks, lane, af[fm]);
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) OB::frag(lb, wn * WTN + fn * 16, ks, lane, bfg[fn]);
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) AR::mma(acc[fm][fn], af[fm], bfg[fn]);
        }
    }

    // Wait until this wave's loads of the block about to be used have landed
    // (at most `ahead` later blocks still in flight).
    template <int A = S - 2>
    __device__ __forceinline__ static void wait_block(int ahead)
    {
        if constexpr (A == 0) {
            wait_vm<0>();
        } else {
            if (ahead >= A) wait_vm<A * PER_WAVE>();
            else wait_block<A - 1>(ahead);
        }
    }
};

// One TxT tile (T = 64 or 128) of a grouped int8 / fp32 instance; `started`
// runs once the first k-blocks' loads are issued, `done_loading` once the
// tile's loads have all landed (the item loop settles the next
// item's prefetched description there).  Starts with a barrier: the
// previous item's LDS readers are done.
template <class AR, int TS, bool AKC, bool BKC, int BM = TS, int NW = 4, class Started, class Hook>
__device__ __forceinline__ void gdma_job(const TileJob& j, uint32_t abytes, uint32_t bbytes, char* smem,
                                         Started&& started, Hook&& done_loading)
{
    constexpr int BK = GDmaCfg<AR, TS>::BK, S = GDmaCfg<AR, TS>::S;
    using T = GDmaTile<AR, AKC, BKC, BM, TS, BK, S, NW>;
    using Out = typename AR::Out;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % T::WAVES_M, wn = wave / T::WAVES_M;
    const int nk = (j.k + BK - 1) / BK;
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(j.A, abytes);
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(j.B, bbytes);
    // byte footprint of a k-block (first to last element it reads), to find
    // the blocks that can hold a chunk crossing the end of a matrix
    constexpr int E = AR::E;
    const long sra = AKC ? j.lda : 1, ska = AKC ? 1 : j.lda, srb = BKC ? j.ldb : 1, skb = BKC ? 1 : j.ldb;
    const uint32_t a_org = (uint32_t)((long)j.i0 * sra * E), b_org = (uint32_t)((long)j.j0 * srb * E);
    const uint32_t a_kstep = (uint32_t)(BK * ska * E), b_kstep = (uint32_t)(BK * skb * E);
    const uint32_t a_span = (uint32_t)(((BM - 1) * sra + (BK - 1) * ska + 1) * E);
    const uint32_t b_span = (uint32_t)(((TS - 1) * srb + (BK - 1) * skb + 1) * E);

    typename AR::Acc acc[T::FM][T::FN];
#pragma unroll
    for (int fm = 0; fm < T::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < T::FN; ++fn) acc[fm][fn] = typename AR::Acc{};

    lds_barrier();
#pragma unroll
    for (int s = 0; s < S - 1; ++s)
        if (s < nk) T::issue(j, ra, rb, s, smem + s * T::STAGE, wave, lane);
    started();
    for (int kt = 0; kt < nk; ++kt) {
        char* cur = smem + (kt % S) * T::STAGE;
        T::wait_block(min(nk - 1 - kt, S - 2));
        // fixes in LDS by the lanes that loaded the chunks, before the
        // barrier that publishes the block
        const int k0 = kt * BK, kv = j.k - k0;
        bool fixed = false;
        if constexpr (AR::E == 1) {
            if (a_org + (uint32_t)kt * a_kstep + a_span > abytes) {
                T::template fix_end<typename T::OA, T::A_PW>(cur, ra, abytes, j.i0, k0, j.lda, wave, lane);
                fixed = true;
            }
            if (b_org + (uint32_t)kt * b_kstep + b_span > bbytes) {
                T::template fix_end<typename T::OB, T::B_PW>(cur + T::OA::BYTES, rb, bbytes, j.j0, k0, j.ldb, wave,
                                                             lane);
                fixed = true;
            }
        }
        if (kv < BK) {
            if constexpr (AKC) T::template clear_k<typename T::OA, T::A_PW>(cur, kv, wave, lane);
            if constexpr (BKC) T::template clear_k<typename T::OB, T::B_PW>(cur + T::OA::BYTES, kv, wave, lane);
            fixed = fixed || AKC || BKC;
        }
        if (fixed) wait_lgkm0();
        lds_barrier();
        if (kt + S - 1 < nk) T::issue(j, ra, rb, kt + S - 1, smem + ((kt + S - 1) % S) * T::STAGE, wave, lane);
        T::compute(acc, cur, wm, wn, lane);
    }
    done_loading();

    Out* C = static_cast<Out*>(j.C);
    const int row0 = j.i0 + wm * T::WTM + 4 * (lane >> 4), col0 = j.j0 + wn * T::WTN + (lane & 15);
    gstore_block<AR, T::FM, T::FN>(C, j.ldc, row0, col0, acc, j.s, j.beta0 != 0, j.m, j.n);
}

// One wave's 32x32 block of a grouped int8 / fp32 instance (a one-wave item
// of a bundle) with the same direct-to-LDS loads: k-blocks of 4 KiB (int8
// 32x64 and 64x32, fp32 32x16 and 16x32), three stages in the wave's own
// 12 KiB of LDS, so two k-blocks are in flight while one is multiplied and
// no staging registers are held.  Only the wave reads its LDS: it waits for
// its own loads and needs no barrier.  `started` runs once the first
// k-blocks' loads are issued.
constexpr int G_WAVE_LDS = 12288;

template <class AR>
struct GWaveCfg {
    static constexpr int BK = AR::ID == ARITH_I8 ? 64 : 16, S = 3;
};

template <class AR, bool AKC, bool BKC, class Started>
__device__ __forceinline__ void gdma_wave_job(const TileJob& j, uint32_t abytes, uint32_t bbytes, char* slot,
                                              Started&& started)
{
    constexpr int BK = GWaveCfg<AR>::BK, S = GWaveCfg<AR>::S, TS = 32;
    using T = GDmaTile<AR, AKC, BKC, TS, TS, BK, S, 1>;
    static_assert(T::LDS_BYTES <= G_WAVE_LDS, "three stages in the wave's LDS");
    using Out = typename AR::Out;
    const int lane = threadIdx.x & 63;
    const int nk = (j.k + BK - 1) / BK;
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(j.A, abytes);
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(j.B, bbytes);
    constexpr int E = AR::E;
    const long sra = AKC ? j.lda : 1, ska = AKC ? 1 : j.lda, srb = BKC ? j.ldb : 1, skb = BKC ? 1 : j.ldb;
    const uint32_t a_org = (uint32_t)((long)j.i0 * sra * E), b_org = (uint32_t)((long)j.j0 * srb * E);
    const uint32_t a_kstep = (uint32_t)(BK * ska * E), b_kstep = (uint32_t)(BK * skb * E);
    const uint32_t a_span = (uint32_t)(((TS - 1) * sra + (BK - 1) * ska + 1) * E);
    const uint32_t b_span = (uint32_t)(((TS - 1) * srb + (BK - 1) * skb + 1) * E);

    typename AR::Acc acc[T::FM][T::FN];
#pragma unroll
    for (int fm = 0; fm < T::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < T::FN; ++fn) acc[fm][fn] = typename AR::Acc{};

    __asm__ __volatile__("" ::: "memory");
#pragma unroll
    for (int s = 0; s < S - 1; ++s)
        if (s < nk) T::issue(j, ra, rb, s, slot + s * T::STAGE, 0, lane);
    started();
    for (int kt = 0; kt < nk; ++kt) {
        char* cur = slot + (kt % S) * T::STAGE;
        T::wait_block(min(nk - 1 - kt, S - 2));
        const int k0 = kt * BK, kv = j.k - k0;
        if constexpr (AR::E == 1) {
            if (a_org + (uint32_t)kt * a_kstep + a_span > abytes)
                T::template fix_end<typename T::OA, T::A_PW>(cur, ra, abytes, j.i0, k0, j.lda, 0, lane);
            if (b_org + (uint32_t)kt * b_kstep + b_span > bbytes)
                T::template fix_end<typename T::OB, T::B_PW>(cur + T::OA::BYTES, rb, bbytes, j.j0, k0, j.ldb, 0,
                                                             lane);
        }
        if (kv < BK) {
            if constexpr (AKC) T::template clear_k<typename T::OA, T::A_PW>(cur, kv, 0, lane);
            if constexpr (BKC) T::template clear_k<typename T::OB, T::B_PW>(cur + T::OA::BYTES, kv, 0, lane);
        }
        // the fixes and the previous block's fragment reads have completed
        // before the block is read and its stage refilled
        wait_lgkm0();
        __asm__ __volatile__("" ::: "memory");
        if (kt + S - 1 < nk) T::issue(j, ra, rb, kt + S - 1, slot + ((kt + S - 1) % S) * T::STAGE, 0, lane);
        T::compute(acc, cur, 0, 0, lane);
    }
    __asm__ __volatile__("" ::: "memory");

    Out* C = static_cast<Out*>(j.C);
    gstore_block<AR, T::FM, T::FN>(C, j.ldc, j.i0 + 4 * (lane >> 4), j.j0 + (lane & 15), acc, j.s, j.beta0 != 0,
                                   j.m, j.n);
}

template <class AR, int TS, int BM = TS, int NW = 4>
constexpr int gdma_lds_bytes()
{
    using T = GDmaTile<AR, true, true, BM, TS, GDmaCfg<AR, TS>::BK, GDmaCfg<AR, TS>::S, NW>;
    using U = GDmaTile<AR, false, false, BM, TS, GDmaCfg<AR, TS>::BK, GDmaCfg<AR, TS>::S, NW>;
    return T::LDS_BYTES > U::LDS_BYTES ? T::LDS_BYTES : U::LDS_BYTES;
}

template <class AR, int TS, class Started, class Hook>
__device__ __forceinline__ void g_run_dma(const TileJob& j, uint32_t abytes, uint32_t bbytes, char* smem,
                                          Started&& started, Hook&& done_loading)
{
    if (j.akc) {
        if (j.bkc) gdma_job<AR, TS, true, true>(j, abytes, bbytes, smem, started, done_loading);
        else gdma_job<AR, TS, true, false>(j, abytes, bbytes, smem, started, done_loading);
    } else {
        if (j.bkc) gdma_job<AR, TS, false, true>(j, abytes, bbytes, smem, started, done_loading);
        else gdma_job<AR, TS, false, false>(j, abytes, bbytes, smem, started, done_loading);
    }
}

// Batched kernel of BM x TS direct-to-LDS tiles of NW waves (configs
// gdma256x128: 8 waves of 64x64; gdma128, gdma64: 4 waves; int8, fp32 and the
// bf16 split forms of fp32): persistent workgroups walk the (instance, tile)
// work ids, neighbouring tiles of an instance on one XCD.  abytes / bbytes: extent of every instance's stored A
// / B.
template <class AR, bool AKC, bool BKC, int TS, int BM = TS, int NW = 4>
__global__ void __launch_bounds__(64 * NW)
    __attribute__((amdgpu_waves_per_eu(AR::NPLANE > 1 ? 2 : NW == 8 ? 4 : 3)))
    gdma_gemm(const KParams p, uint32_t abytes, uint32_t bbytes)
{
    __shared__ __attribute__((aligned(1024))) char smem[gdma_lds_bytes<AR, TS, BM, NW>()];
    const int tiles_per_inst = p.tiles_m * p.tiles_n;
    for (long t = blockIdx.x; t < p.total; t += gridDim.x) {
        const long w = xcd_remap(t, p.total);
        const int inst = (int)(w / tiles_per_inst);
        const int rem = (int)(w % tiles_per_inst);
        const TileJob j = make_job(p, inst, (rem % p.tiles_m) * BM, (rem / p.tiles_m) * TS);
        gdma_job<AR, TS, AKC, BKC, BM, NW>(j, abytes, bbytes, smem, [] {}, [] {});
    }
}

}  // namespace ssb
