// grouped_kernel.hpp -- the grouped (variable-size, mixed-type) batched GEMM:
// one persistent kernel per family walks a device list of entries built on
// the host (grouped_host.cpp).
//
// Entry = one output tile of one instance (32x32, 64x64 or 128x128: the
// shape chosen for the instance, or a smaller one for the rows / columns
// left over past its last whole tile), or a bundle of four one-wave 32x32
// items; through the instance come its group's arithmetic form, operand
// layouts and scalars.  Entries sit in per-XCD queues, instance by instance,
// and workgroups claim them (below), so long tiles start first and short
// ones fill in behind.  A launch writes C and the plan's claim counters,
// which its last workgroup returns to zero (a captured graph replays it as
// is).
//
// Operands are read with buffer loads through a descriptor that spans the
// stored matrix: whole 16-byte (or 4-byte) words at any leading dimension
// and alignment, where a dword past the end of the matrix reads as zero
// instead of faulting.  Partial tiles take the same loads as whole ones;
// only a k tail (and, for int8, the word holding a matrix's last bytes) is
// fixed, in registers (32x32 tiles; the fp64 and bf16 one-wave items) or in
// LDS (the direct-to-LDS int8 / fp32 tiles and one-wave items of
// grouped_dma.hpp).
//
// Included by one .hip file per family.
#pragma once
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "gemm_kernels.hpp"
#include "grouped.hpp"
#include "grouped_dma.hpp"
#include "launch.hpp"
#include "occupancy.hpp"

namespace ssb {

// Buffer-load staging of one operand tile (ROWS x BK), same unit and word
// layout as StagerImpl<KC>, so StagerImpl::store writes it to LDS.
//
// Every word is read whole through the descriptor, which spans the stored
// matrix (`ext` bytes): a dword past the end reads as zero.  Rows (columns)
// outside the instance are not masked: they only feed rows (columns) of the
// product that are never stored.  What has to be fixed:
//  * a K-contiguous operand in a k-block that reaches past K: a word's
//    elements at k >= K hold padding or the next row and are cleared
//    (`clear_k`); an M/N-contiguous operand's words at k >= K start past the
//    end of the matrix and read as zero;
//  * int8: the dword holding the matrix's last bytes can reach past the end
//    and read as zero as well, so a word that crosses the end is read again
//    as the WB bytes ending exactly at `ext` and shifted into place
//    (`fix_end`; a wave takes that branch only when one of its lanes holds
//    such a word, which only the tile holding the last element of the
//    matrix, in its last k-block, can have).  A matrix smaller than one word
//    is read element by element there.
template <class AR, bool KC, int ROWS, int BK, int NT>
struct GLoad {
    using I = StagerImpl<AR, KC, ROWS, BK, NT>;
    static constexpr int WB = I::WB, E = AR::E;

    __device__ __forceinline__ static bool crossing(uint32_t off, uint32_t ext)
    {
        return off < ext && off + WB > ext;
    }

    __device__ __forceinline__ static void read_word(uint32_t* d, __amdgpu_buffer_rsrc_t rs, uint32_t off)
    {
        if constexpr (WB == 16) {
            const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
            d[0] = v.x;
            d[1] = v.y;
            d[2] = v.z;
            d[3] = v.w;
        } else {
            d[0] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0);
        }
    }

    // The elements of the word at `off` that end inside the matrix, one load each.
    __device__ __forceinline__ static void read_elems(uint32_t* d, __amdgpu_buffer_rsrc_t rs, uint32_t off,
                                                      uint32_t ext)
    {
#pragma unroll
        for (int i = 0; i < WB / 4; ++i) d[i] = 0u;
#pragma unroll
        for (int e = 0; e < WB / E; ++e) {
            const uint32_t p = off + e * E;
            if (p + E > ext) break;
            if constexpr (E == 1) {
                const uint32_t b = __builtin_amdgcn_raw_buffer_load_b8(rs, (int)p, 0, 0);
                d[e >> 2] |= b << (8 * (e & 3));
            } else if constexpr (E == 4) {
                d[e] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)p, 0, 0);
            } else {
                const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)p, 0, 0);
                d[2 * e] = v.x;
                d[2 * e + 1] = v.y;
            }
        }
    }

    // off0: byte offset of the tile origin (row 0, k 0 of the k-block).
    __device__ __forceinline__ static void load(uint32_t* raw, __amdgpu_buffer_rsrc_t rs, uint32_t off0, long sr,
                                                long sk, int tid)
    {
#pragma unroll
        for (int x = 0; x < I::UPT; ++x) {
            const int u = tid + x * NT;
            if (!I::active(u)) break;
            int r0, k0;
            I::unit_coords(u, r0, k0);
#pragma unroll
            for (int w = 0; w < I::NW; ++w)
                read_word(raw + x * I::UD + w * I::WD, rs, off0 + (uint32_t)I::word_off(r0, k0, w, sr, sk));
        }
    }

    // Words that cross the end of the matrix, read again and shifted.
    __device__ __forceinline__ static void fix_end(uint32_t* raw, __amdgpu_buffer_rsrc_t rs, uint32_t off0,
                                                   uint32_t ext, long sr, long sk, int tid)
    {
        bool any = false;
#pragma unroll
        for (int x = 0; x < I::UPT; ++x) {
            const int u = tid + x * NT;
            if (!I::active(u)) break;
            int r0, k0;
            I::unit_coords(u, r0, k0);
#pragma unroll
            for (int w = 0; w < I::NW; ++w) any |= crossing(off0 + (uint32_t)I::word_off(r0, k0, w, sr, sk), ext);
        }
        if (!__any(any)) return;
#pragma unroll
        for (int x = 0; x < I::UPT; ++x) {
            const int u = tid + x * NT;
            if (!I::active(u)) break;
            int r0, k0;
            I::unit_coords(u, r0, k0);
#pragma unroll
            for (int w = 0; w < I::NW; ++w) {
                uint32_t* d = raw + x * I::UD + w * I::WD;
                const uint32_t off = off0 + (uint32_t)I::word_off(r0, k0, w, sr, sk);
                if (!crossing(off, ext)) continue;
                if (ext < (uint32_t)WB) {
                    read_elems(d, rs, off, ext);
                    continue;
                }
                read_word(d, rs, ext - WB);
                // byte j of the word is byte j + sh of what was read
                const uint32_t sh = off - (ext - WB);
                if constexpr (WB == 16) {
                    const uint32_t q = sh >> 2, r = sh & 3;
                    uint32_t t[5];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t a = i + q;
                        t[i] = a == 0 ? d[0] : a == 1 ? d[1] : a == 2 ? d[2] : a == 3 ? d[3] : 0u;
                    }
                    t[4] = 0u;
#pragma unroll
                    for (int i = 0; i < 4; ++i) d[i] = __builtin_amdgcn_alignbyte(t[i + 1], t[i], r);
                } else {
                    d[0] = __builtin_amdgcn_alignbyte(0u, d[0], sh);
                }
            }
        }
    }

    // K-contiguous operand: clear the elements at tile-relative k >= k_valid.
    __device__ __forceinline__ static void clear_k(uint32_t* raw, int k_valid, int tid)
    {
        static_assert(KC, "M/N-contiguous words past K read as zero");
#pragma unroll
        for (int x = 0; x < I::UPT; ++x) {
            const int u = tid + x * NT;
            if (!I::active(u)) break;
            int r0, k0;
            I::unit_coords(u, r0, k0);
            clear_tail<WB, E>(raw + x * I::UD, min(max(k_valid - k0, 0), (int)I::WE));
        }
    }
};

// One operand's staging registers with its layout chosen at run time.
template <class AR, int ROWS, int BK, int NT>
struct GStage {
    using LK = GLoad<AR, true, ROWS, BK, NT>;
    using LN = GLoad<AR, false, ROWS, BK, NT>;
    static constexpr int RAWD = LK::I::RAWD > LN::I::RAWD ? LK::I::RAWD : LN::I::RAWD;
    uint32_t raw[RAWD];

    __device__ __forceinline__ void load(bool kc, __amdgpu_buffer_rsrc_t rs, uint32_t off0, long sr, long sk, int tid)
    {
        if (kc) LK::load(raw, rs, off0, sr, sk, tid);
        else LN::load(raw, rs, off0, sr, sk, tid);
    }
    __device__ __forceinline__ void fix_end(bool kc, __amdgpu_buffer_rsrc_t rs, uint32_t off0, uint32_t ext, long sr,
                                            long sk, int tid)
    {
        if (kc) LK::fix_end(raw, rs, off0, ext, sr, sk, tid);
        else LN::fix_end(raw, rs, off0, ext, sr, sk, tid);
    }
    __device__ __forceinline__ void clear_k(bool kc, int k_valid, int tid)
    {
        if (kc) LK::clear_k(raw, k_valid, tid);
    }
    __device__ __forceinline__ void store(bool kc, char* lds, int plane_bytes, int rowb, int tid) const
    {
        if (kc) LK::I::store(raw, lds, plane_bytes, rowb, tid);
        else LN::I::store(raw, lds, plane_bytes, rowb, tid);
    }
};

// One BM x BN tile of one grouped instance, 4 waves (2 x 2), LDS
// double-buffered, global loads of k-block kt+1 in flight while kt is
// multiplied.  abytes / bbytes: extent of the stored A / B in bytes.
// `started` runs once the first k-block's loads are issued, `first_block`
// once that block is in LDS (the entry loop's claim and prefetch hooks).
template <class AR, int BM, int BN, int BK, class Started, class Hook>
__device__ __forceinline__ void gtile_job(const TileJob& j, uint32_t abytes, uint32_t bbytes, char* smem,
                                          Started&& started, Hook&& first_block)
{
    using S = TileShape<AR, true, true, BM, BN, BK, 2, 2>;
    using Out = typename AR::Out;
    constexpr int NT = S::NT;
    constexpr int E = AR::E;
    using SA = GStage<AR, BM, BK, NT>;
    using SB = GStage<AR, BN, BK, NT>;
    const bool AKC = j.akc != 0, BKC = j.bkc != 0;
    // Opaque to the optimiser: keeps the lane-dependent LDS and global
    // offsets of this tile shape from being hoisted out of the item loop,
    // where they would stay live beside those of every other shape.
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    const long sra = AKC ? j.lda : 1, ska = AKC ? 1 : j.lda;
    const long srb = BKC ? j.ldb : 1, skb = BKC ? 1 : j.ldb;
    const int nk = (j.k + BK - 1) / BK;
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(j.A, abytes);
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(j.B, bbytes);
    const uint32_t a0 = (uint32_t)((long)j.i0 * sra * E), b0 = (uint32_t)((long)j.j0 * srb * E);
    const uint32_t a_kstep = (uint32_t)(BK * ska * E), b_kstep = (uint32_t)(BK * skb * E);

    SA sa;
    SB sb;
    typename AR::Acc acc[S::FM][S::FN];
#pragma unroll
    for (int fm = 0; fm < S::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < S::FN; ++fn) acc[fm][fn] = typename AR::Acc{};

    auto load_step = [&](int kt) {
        sa.load(AKC, ra, a0 + (uint32_t)kt * a_kstep, sra, ska, tid);
        sb.load(BKC, rb, b0 + (uint32_t)kt * b_kstep, srb, skb, tid);
    };
    // int8 words crossing the end of a matrix: only where the k-block's
    // footprint (first to last element it reads) reaches past the end; then
    // the k tail of the last k-block
    const uint32_t a_span = (uint32_t)(((BM - 1) * sra + (BK - 1) * ska + 1) * E);
    const uint32_t b_span = (uint32_t)(((BN - 1) * srb + (BK - 1) * skb + 1) * E);
    auto store_step = [&](int kt, int buf) {
        if constexpr (E == 1) {
            const uint32_t ao = a0 + (uint32_t)kt * a_kstep, bo = b0 + (uint32_t)kt * b_kstep;
            if (ao + a_span > abytes) sa.fix_end(AKC, ra, ao, abytes, sra, ska, tid);
            if (bo + b_span > bbytes) sb.fix_end(BKC, rb, bo, bbytes, srb, skb, tid);
        }
        if (kt == nk - 1) {
            const int kv = j.k - kt * BK;
            if (kv < BK) {
                sa.clear_k(AKC, kv, tid);
                sb.clear_k(BKC, kv, tid);
            }
        }
        char* base = smem + buf * S::BUF;
        sa.store(AKC, base, S::A_PLANE, S::ROWB, tid);
        sb.store(BKC, base + AR::NPLANE * S::A_PLANE, S::B_PLANE, S::ROWB, tid);
    };

    load_step(0);
    started();
    __syncthreads();  // previous user of smem is done
    store_step(0, 0);
    __syncthreads();
    first_block();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_step(kt + 1);
        const char* la = smem + buf * S::BUF;
        tile_compute<AR, S, BK>(acc, la, la + AR::NPLANE * S::A_PLANE, wm, wn, lane);
        if (kt + 1 < nk) store_step(kt + 1, buf ^ 1);
        __syncthreads();
    }

    Out* C = static_cast<Out*>(j.C);
    const int row0 = j.i0 + wm * S::WTM + 4 * (lane >> 4), col0 = j.j0 + wn * S::WTN + (lane & 15);
    gstore_block<AR, S::FM, S::FN>(C, j.ldc, row0, col0, acc, j.s, j.beta0 != 0, j.m, j.n);
}

template <class AR, int CFG>
struct GShape {
    static constexpr int BM = CFG == GCFG_32 ? 32 : CFG == GCFG_64 ? 64 : 128;
    static constexpr int BN = BM;
    static constexpr int BK = grouped_bk(AR::ID, CFG);
    static constexpr int LDS = TileShape<AR, true, true, BM, BN, BK, 2, 2>::LDS_BYTES;
};

// The int8 and fp32 forms run their 64x64 and 128x128 tiles with
// direct-to-LDS stages, the other forms register-staged double buffers.
template <class AR>
constexpr bool g_dma()
{
    return AR::ID == ARITH_I8 || AR::ID == ARITH_F32;
}
template <class AR, int CFG>
constexpr int g_tile_lds()
{
    if constexpr (g_dma<AR>() && CFG != GCFG_32) return gdma_lds_bytes<AR, CFG == GCFG_64 ? 64 : 128>();
    else return GShape<AR, CFG>::LDS;
}

template <class AR>
constexpr int g_lds_max()
{
    constexpr int a = g_tile_lds<AR, GCFG_32>(), b = g_tile_lds<AR, GCFG_64>(), c = g_tile_lds<AR, GCFG_128>();
    constexpr int ab = a > b ? a : b;
    constexpr int t = grouped_max_cfg(AR::ID) == GCFG_128 ? (ab > c ? ab : c) : ab;
    // four waves' one-wave items side by side
    constexpr int w = AR::ID == ARITH_I8 || AR::ID == ARITH_F32 ? 4 * G_WAVE_LDS : 0;
    return t > w ? t : w;
}

template <class AR, class Started, class Hook>
__device__ __forceinline__ void g_run(int cfg, const TileJob& j, uint32_t abytes, uint32_t bbytes, char* smem,
                                      Started&& started, Hook&& first_block)
{
    if (cfg == GCFG_32) {
        using G = GShape<AR, GCFG_32>;
        gtile_job<AR, G::BM, G::BN, G::BK>(j, abytes, bbytes, smem, started, first_block);
        return;
    }
    if constexpr (grouped_max_cfg(AR::ID) == GCFG_128) {
        if (cfg == GCFG_128) {
            if constexpr (g_dma<AR>()) {
                g_run_dma<AR, 128>(j, abytes, bbytes, smem, started, first_block);
            } else {
                using G = GShape<AR, GCFG_128>;
                gtile_job<AR, G::BM, G::BN, G::BK>(j, abytes, bbytes, smem, started, first_block);
            }
            return;
        }
    }
    if constexpr (g_dma<AR>()) {
        g_run_dma<AR, 64>(j, abytes, bbytes, smem, started, first_block);
    } else {
        using G = GShape<AR, GCFG_64>;
        gtile_job<AR, G::BM, G::BN, G::BK>(j, abytes, bbytes, smem, started, first_block);
    }
}

// One wave's 32x32 block of a grouped instance (a one-wave item of a
// bundle), operand layouts chosen at run time: int8 and fp32 through the
// wave's direct-to-LDS job in `slot` (G_WAVE_LDS bytes of LDS private to the
// wave), fp64 and the bf16 split forms with operands loaded straight into
// registers.  `started` runs once the item's first loads are issued.
template <class AR, class Started>
__device__ __forceinline__ void g_run_wave(const TileJob& j, uint32_t abytes, uint32_t bbytes, char* slot,
                                           Started&& started)
{
    if constexpr (AR::ID == ARITH_I8 || AR::ID == ARITH_F32) {
        if (j.akc) {
            if (j.bkc) gdma_wave_job<AR, true, true>(j, abytes, bbytes, slot, started);
            else gdma_wave_job<AR, true, false>(j, abytes, bbytes, slot, started);
        } else {
            if (j.bkc) gdma_wave_job<AR, false, true>(j, abytes, bbytes, slot, started);
            else gdma_wave_job<AR, false, false>(j, abytes, bbytes, slot, started);
        }
    } else {
        started();
        if (j.akc) {
            if (j.bkc) direct_job<AR, true, true, 32, 32>(j);
            else direct_job<AR, true, false, 32, 32>(j);
        } else {
            if (j.bkc) direct_job<AR, false, true, 32, 32>(j);
            else direct_job<AR, false, false, 32, 32>(j);
        }
    }
}

// An item's description travels lane-packed in one VGPR: lanes 0..15 hold
// the 16 words of its record, lanes 16..21 its A, B and C pointers (low word
// first), lanes 32..47 the record of the item after it.  One vector load per
// item fills it; v_readlane decodes it into scalars.
__device__ __forceinline__ uint32_t g_fetch(const GroupedDevice& d, int it, int inst, int it_after, int lane)
{
    const int w = min(max(lane - 16, 0), 5);
    const void* const* arr = w < 2 ? d.A : w < 4 ? d.B : const_cast<const void* const*>(d.C);
    const uint32_t* ptr_word = reinterpret_cast<const uint32_t*>(arr + inst) + (w & 1);
    const int r = (lane >= 32 && lane < 48) ? it_after : it;
    const uint32_t* rec_word = reinterpret_cast<const uint32_t*>(d.recs + r) + (lane & (GREC_WORDS - 1));
    return *gbl((lane >= 16 && lane < 22) ? ptr_word : rec_word);
}
// A copy of v made once v's load has landed: a value carried to the next
// item without a load in flight behind it, so using it there waits for
// nothing (otherwise the wait, in order behind this item's stores, would
// also wait for them).
__device__ __forceinline__ uint32_t g_settle(uint32_t v)
{
    uint32_t r;
    __asm__ volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

struct GWork {
    TileJob j;
    int cfg, form;
    uint32_t abytes, bbytes;
};

__device__ __forceinline__ GWork g_decode(uint32_t v)
{
    auto w = [&](int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); };
    auto ptr = [&](int l) { return (uintptr_t)(((uint64_t)w(l + 1) << 32) | w(l)); };
    GWork g;
    g.j.A = reinterpret_cast<const char*>(ptr(16));
    g.j.B = reinterpret_cast<const char*>(ptr(18));
    g.j.C = reinterpret_cast<void*>(ptr(20));
    g.j.i0 = (int)w(1);
    g.j.j0 = (int)w(2);
    const uint32_t bits = w(3);
    g.j.m = (int)w(4);
    g.j.n = (int)w(5);
    g.j.k = (int)w(6);
    g.j.lda = (int)w(7);
    g.j.ldb = (int)w(8);
    g.j.ldc = (int)w(9);
    g.abytes = w(10);
    g.bbytes = w(11);
    const uint32_t a0 = w(12), a1 = w(13), b0 = w(14), b1 = w(15);
    g.j.s.af = __builtin_bit_cast(float, a0);
    g.j.s.bf = __builtin_bit_cast(float, b0);
    g.j.s.ad = __builtin_bit_cast(double, ((uint64_t)a1 << 32) | a0);
    g.j.s.bd = __builtin_bit_cast(double, ((uint64_t)b1 << 32) | b0);
    g.j.s.ai = (int32_t)a0;
    g.j.s.bi = (int32_t)b0;
    g.cfg = (int)(bits & 3u);
    g.form = (int)((bits >> 2) & 7u);
    g.j.akc = (int)((bits >> 5) & 1u);
    g.j.bkc = (int)((bits >> 6) & 1u);
    g.j.beta0 = (int)((bits >> 7) & 1u);
    return g;
}

// Entries are claimed from the per-XCD queues by the workgroup's leader
// thread, one atomic per entry on the queue's counter.  Workgroup b's home
// queue is b % GQ (b runs on XCD b % 8, whose L2 then holds the queue's
// instances); the n workgroups of a queue take its first 3 n entries at
// fixed places (workgroup r: entries r, r + n, r + 2 n, so a launch does not
// start with a burst of atomics on one counter) and claim the rest one at a
// time, moving on to the next queues when theirs runs dry, so the queues
// drain together whatever the entries actually cost (a fixed assignment had
// left some workgroups running 30 % longer than the median on the same
// estimated work).  An atomic's return retires in order with the wave's
// other memory accesses, so a claim is issued right after an entry's first
// loads and examined once all of them have landed (g_claim_finish).
struct GClaim {
    int q;      // queue drawn from
    int tried;  // queues found empty (queues only empty further)
};
// The number of workgroups whose home queue is q, of G dealt b % GQ.
__device__ __forceinline__ int g_home_count(int G, int q)
{
    return G > q ? (G - q + GQ - 1) / GQ : 0;
}
__device__ __forceinline__ int g_claim_issue(int* ctr, const GClaim& c)
{
    return c.tried >= GQ ? INT_MAX : atomicAdd(ctr + c.q * GCTR_STRIDE, 1);
}
// The entry the claim `r` on queue c.q gives (past the queue's fixed
// entries), moving on to the next queues while they are empty; -1 when
// every queue is.
__device__ __forceinline__ int g_claim_finish(const int* off, int* ctr, int G, GClaim& c, int r)
{
    while (c.tried < GQ) {
        const int x = 3 * g_home_count(G, c.q) + r;
        if (x < off[c.q + 1] - off[c.q]) return off[c.q] + x;
        c.q = c.q + 1 == GQ ? 0 : c.q + 1;
        if (++c.tried >= GQ) break;
        r = atomicAdd(ctr + c.q * GCTR_STRIDE, 1);
    }
    return -1;
}
// Fixed entry s (0, 1, 2) of workgroup `rank` of queue q (n workgroups), -1 past the queue.
__device__ __forceinline__ int g_fixed(const int* off, int q, int n, int rank, int s)
{
    const int x = rank + s * n;
    return x < off[q + 1] - off[q] ? off[q] + x : -1;
}

// Waves per SIMD the family's kernel is compiled for: three workgroups per
// CU (48 KiB of LDS each) where the registers allow; the bf16 split forms
// hold two or three operand planes per fragment and keep two.
constexpr int g_waves_per_eu(int second_form)
{
    return second_form == ARITH_BF16X3 || second_form == ARITH_BF16X6 ? 2 : 3;
}

// FA: the family's first form; FB: its second (== FA for one-form families).
// Each workgroup runs the entries it claims: a tile with all four waves, or
// a bundle of four one-wave items, one per wave (tiles and bundles of a
// queue sorted together by cost, so small items run beside the tiles).
// Claims run three entries ahead, and while an entry runs each wave reads
// the record of its part of the entry after next and the operand pointers
// of the next one, so an entry starts with its description at hand instead
// of a chain of dependent loads, each in order behind the previous entry's
// stores.  The last workgroup to finish re-arms the counters for the next
// launch (a captured graph replays the kernel as it is); launches of one
// plan must not overlap.
template <class FA, class FB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(g_waves_per_eu(FB::ID)))) grouped_gemm(const GroupedDevice d)
{
    constexpr int LA = g_lds_max<FA>(), LB = g_lds_max<FB>();
    __shared__ __attribute__((aligned(16))) char smem[LA > LB ? LA : LB];
    __shared__ int s_claim[3];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const bool leader = threadIdx.x == 0;
    int* const qctr = d.ctrs;
    const int G = (int)gridDim.x, home = (int)blockIdx.x % GQ, rank = (int)blockIdx.x / GQ;
    const int nh = g_home_count(G, home);
    GClaim cs{home, 0};
    int i0 = g_fixed(d.qoff, home, nh, rank, 0), i1 = g_fixed(d.qoff, home, nh, rank, 1),
        i2 = g_fixed(d.qoff, home, nh, rank, 2);
    if (i2 < 0) {
        // no claims left in the home queue for this workgroup: claim at once
        // (the fixed entries that exist come first, in order)
        if (leader) {
            int got[3] = {i0, i1, i2};
#pragma unroll
            for (int x = 0; x < 3; ++x)
                if (got[x] < 0) got[x] = g_claim_finish(d.qoff, qctr, G, cs, g_claim_issue(qctr, cs));
#pragma unroll
            for (int x = 0; x < 3; ++x) s_claim[x] = got[x];
        }
        __syncthreads();
        i0 = __builtin_amdgcn_readfirstlane(s_claim[0]);
        i1 = __builtin_amdgcn_readfirstlane(s_claim[1]);
        i2 = __builtin_amdgcn_readfirstlane(s_claim[2]);
        __syncthreads();
    }
    if (i0 >= 0) {
        // this wave's record of entry e (a missing entry: the current one's)
        auto rec = [&](int e) { return 4 * (e >= 0 ? e : i0) + wave; };
        // the first entry's record, then its pointers with the next record
        const uint32_t r0 = g_fetch(d, rec(i0), 0, rec(i0), lane);
        uint32_t v = g_settle(g_fetch(d, rec(i0), __builtin_amdgcn_readlane((int)r0, 0), rec(i1), lane));
        int slot = 0;
        while (i0 >= 0) {
            const GWork w = g_decode(v);
            const uint32_t t = g_fetch(d, rec(i1), __builtin_amdgcn_readlane((int)v, 32), rec(i2), lane);
            // the claim of the entry three ahead: issued once the entry's first
            // loads are, examined once they have all landed (a claim left in
            // flight across the loop's back edge had the compiler wait for
            // every access in flight, stores included, at the loop's head)
            uint32_t v_next = 0u;
            int cl = INT_MAX;
            auto started = [&]() {
                if (leader) cl = g_claim_issue(qctr, cs);
            };
            auto loaded = [&]() {
                v_next = g_settle(t);
                if (leader) s_claim[slot] = g_claim_finish(d.qoff, qctr, G, cs, cl);
            };
            if (w.cfg == GCFG_W32) {
                if (w.j.m > 0) {
                    char* slotp = smem + G_WAVE_LDS * wave;
                    if (w.form == FA::ID) g_run_wave<FA>(w.j, w.abytes, w.bbytes, slotp, started);
                    else g_run_wave<FB>(w.j, w.abytes, w.bbytes, slotp, started);
                } else {
                    started();
                }
                loaded();
            } else {
                if (w.form == FA::ID) g_run<FA>(w.cfg, w.j, w.abytes, w.bbytes, smem, started, loaded);
                else g_run<FB>(w.cfg, w.j, w.abytes, w.bbytes, smem, started, loaded);
            }
            // publishes the claim (the leader's LDS write has landed) and tells
            // every wave that the entry's LDS readers are done; not
            // __syncthreads, whose fence would also wait for the entry's stores
            wait_lgkm0();
            lds_barrier();
            // entry indices kept in scalar registers: a vector copy at the
            // loop's end had the compiler wait for every memory access in
            // flight (the entry's stores) before it
            const int i3 = __builtin_amdgcn_readfirstlane(s_claim[slot]);
            slot ^= 1;
            v = v_next;
            i0 = i1;
            i1 = i2;
            i2 = i3;
        }
    }
    // every claim of the workgroup has returned; then the last workgroup
    // re-arms the counters
    wait_vm<0>();
    __syncthreads();
    if (leader && atomicAdd(d.ctrs + GCTR_DONE * GCTR_STRIDE, 1) == G - 1) {
        for (int x = 0; x < GCTR_COUNT; ++x)
            __hip_atomic_store(d.ctrs + x * GCTR_STRIDE, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <class FA, class FB>
hipError_t run_grouped(const GroupedDevice& d, hipStream_t st)
{
    if (d.nitems <= 0) return hipSuccess;
    auto kern = grouped_gemm<FA, FB>;
    long slots = (long)kernel_occupancy(kern, 256) * device_cu_count();
    // SSBLAS_MAX_GRID=<n> caps the grid (tests: few workgroups, some queues
    // with no workgroup of their own, every entry past the first claimed)
    if (const char* e = getenv("SSBLAS_MAX_GRID")) {
        const long cap = atol(e);
        if (cap >= 1 && cap < slots) slots = cap;
    }
    const long want = (long)d.nitems;
    const long grid = want < slots ? want : slots;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, d);
    return hipGetLastError();
}

}  // namespace ssb
