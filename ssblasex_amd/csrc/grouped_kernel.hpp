// grouped_kernel.hpp -- the grouped (variable-size, mixed-type) batched GEMM:
// one persistent kernel per family walks a device work list built on the
// host (grouped_host.cpp).
//
// Work item = one output tile of one instance: the tile shape chosen for the
// instance (32x32, 64x64 or 128x128), and through the instance its group's
// arithmetic form, operand layouts and scalars.  Items sit in per-XCD
// queues sorted by estimated cost, longest first, and every workgroup takes
// a fixed share of its queue in turn with the others, so long tiles start
// first and short ones fill in behind.  A launch writes nothing but C (a
// captured graph replays it as is).
//
// Operands are read with buffer loads through a descriptor that spans the
// stored matrix: whole 16-byte (or 4-byte) words at any leading dimension
// and alignment, where a dword past the end of the matrix reads as zero
// instead of faulting.  Elements outside the tile's rows or k range are then
// cleared in registers before the LDS store, so partial tiles take the same
// path as whole ones.
//
// Included by one .hip file per family.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_kernels.hpp"
#include "grouped.hpp"
#include "launch.hpp"
#include "occupancy.hpp"

namespace ssb {

// Elements from index `nvalid` on of a word (WB bytes, E-byte elements) set to 0.
template <int WB, int E>
__device__ __forceinline__ void clear_tail(uint32_t* d, int nvalid)
{
    const int vb = nvalid * E;
#pragma unroll
    for (int i = 0; i < WB / 4; ++i) {
        const int lo = 4 * i;
        const uint32_t m = vb >= lo + 4 ? 0xffffffffu : (vb <= lo ? 0u : ((1u << (8 * (vb - lo))) - 1u));
        d[i] &= m;
    }
}

// Buffer-load staging of one operand tile (ROWS x BK), same unit and word
// layout as StagerImpl<KC>, so StagerImpl::store writes it to LDS.
//
// Whole tiles read plain words (every word lies inside the matrix).  In a
// partial tile a word may cross the end of the stored matrix (`ext` bytes);
// the descriptor's range check drops a dword that reaches past the end, so
// such a word is read instead as the WB bytes that end exactly at `ext` and
// shifted into place by `fix`.  A matrix smaller than one word is read
// element by element.
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
    __device__ __forceinline__ static void load(uint32_t* raw, __amdgpu_buffer_rsrc_t rs, uint32_t off0, uint32_t ext,
                                                long sr, long sk, bool partial, int tid)
    {
#pragma unroll
        for (int x = 0; x < I::UPT; ++x) {
            const int u = tid + x * NT;
            if (!I::active(u)) break;
            int r0, k0;
            I::unit_coords(u, r0, k0);
#pragma unroll
            for (int w = 0; w < I::NW; ++w) {
                const uint32_t off = off0 + (uint32_t)I::word_off(r0, k0, w, sr, sk);
                uint32_t* d = raw + x * I::UD + w * I::WD;
                if (!partial || !crossing(off, ext)) read_word(d, rs, off);
                else if (ext >= (uint32_t)WB) read_word(d, rs, ext - WB);
                else read_elems(d, rs, off, ext);
            }
        }
    }

    // Partial tiles: shift re-read words into place, then clear what lies
    // outside rows_valid x k_valid (tile-relative).
    __device__ __forceinline__ static void fix(uint32_t* raw, uint32_t off0, uint32_t ext, long sr, long sk,
                                               int rows_valid, int k_valid, int tid)
    {
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
                if (crossing(off, ext) && ext >= (uint32_t)WB) {
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
                int nv;
                if constexpr (KC) nv = r0 < rows_valid ? min(max(k_valid - k0, 0), (int)I::WE) : 0;
                else nv = k0 + w < k_valid ? min(max(rows_valid - r0, 0), (int)I::WE) : 0;
                clear_tail<WB, E>(d, nv);
            }
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

    __device__ __forceinline__ void load(bool kc, __amdgpu_buffer_rsrc_t rs, uint32_t off0, uint32_t ext, long sr,
                                         long sk, bool partial, int tid)
    {
        if (kc) LK::load(raw, rs, off0, ext, sr, sk, partial, tid);
        else LN::load(raw, rs, off0, ext, sr, sk, partial, tid);
    }
    __device__ __forceinline__ void fix(bool kc, uint32_t off0, uint32_t ext, long sr, long sk, int rows_valid,
                                        int k_valid, int tid)
    {
        if (kc) LK::fix(raw, off0, ext, sr, sk, rows_valid, k_valid, tid);
        else LN::fix(raw, off0, ext, sr, sk, rows_valid, k_valid, tid);
    }
    __device__ __forceinline__ void store(bool kc, char* lds, int plane_bytes, int rowb, int tid) const
    {
        if (kc) LK::I::store(raw, lds, plane_bytes, rowb, tid);
        else LN::I::store(raw, lds, plane_bytes, rowb, tid);
    }
};

template <int AL>
struct ElemAligned16;
template <>
struct ElemAligned16<4> {
    typedef uint32_t T __attribute__((ext_vector_type(4), aligned(4)));
};
template <>
struct ElemAligned16<8> {
    typedef uint32_t T __attribute__((ext_vector_type(4), aligned(8)));
};

// Fused alpha/beta epilogue of one 16x16 fragment at any ldc and alignment:
// the lane's 4 rows go out as one (unaligned) vector when all are in range.
template <class AR>
__device__ __forceinline__ void gstore_frag(typename AR::Out* C, long ldc, int r0, int c,
                                            const typename AR::Acc& acc, const Scalars& sc, bool beta0, int m,
                                            int n)
{
    using Out = typename AR::Out;
    // 16 bytes at element (not vector) alignment
    using u32x4_e = typename ElemAligned16<sizeof(Out)>::T;
    constexpr int NV = 4 * (int)sizeof(Out) / 16;
    if (c >= n) return;
    Out* dst = C + (long)c * ldc + r0;
    if (r0 + 4 <= m) {
        u32x4_e w[NV];
        Out old[4] = {Out(0), Out(0), Out(0), Out(0)};
        if (!beta0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) w[i] = *gbl(reinterpret_cast<const u32x4_e*>(dst) + i);
            __builtin_memcpy(old, w, sizeof(old));
        }
        Out res[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) res[v] = epi(acc[v], old[v], sc, beta0);
        __builtin_memcpy(w, res, sizeof(res));
#pragma unroll
        for (int i = 0; i < NV; ++i) *gbl(reinterpret_cast<u32x4_e*>(dst) + i) = w[i];
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (r0 + v < m) {
                const Out o = beta0 ? Out(0) : *gbl(dst + v);
                *gbl(dst + v) = epi(acc[v], o, sc, beta0);
            }
        }
    }
}

// One BM x BN tile of one grouped instance, 4 waves (2 x 2), LDS
// double-buffered, global loads of k-block kt+1 in flight while kt is
// multiplied.  abytes / bbytes: extent of the stored A / B in bytes.
template <class AR, int BM, int BN, int BK>
__device__ __forceinline__ void gtile_job(const TileJob& j, uint32_t abytes, uint32_t bbytes, char* smem)
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
    const int rows_a = min(BM, j.m - j.i0);
    const int rows_b = min(BN, j.n - j.j0);
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

    // a k-block of a tile with rows or k outside the matrix is partial
    auto load_step = [&](int kt) {
        const bool kpart = j.k - kt * BK < BK;
        sa.load(AKC, ra, a0 + (uint32_t)kt * a_kstep, abytes, sra, ska, kpart || rows_a < BM, tid);
        sb.load(BKC, rb, b0 + (uint32_t)kt * b_kstep, bbytes, srb, skb, kpart || rows_b < BN, tid);
    };
    auto store_step = [&](int kt, int buf) {
        const int kv = j.k - kt * BK;
        if (rows_a < BM || kv < BK) sa.fix(AKC, a0 + (uint32_t)kt * a_kstep, abytes, sra, ska, rows_a, kv, tid);
        if (rows_b < BN || kv < BK) sb.fix(BKC, b0 + (uint32_t)kt * b_kstep, bbytes, srb, skb, rows_b, kv, tid);
        char* base = smem + buf * S::BUF;
        sa.store(AKC, base, S::A_PLANE, S::ROWB, tid);
        sb.store(BKC, base + AR::NPLANE * S::A_PLANE, S::B_PLANE, S::ROWB, tid);
    };

    load_step(0);
    __syncthreads();  // previous user of smem is done
    store_step(0, 0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_step(kt + 1);
        const char* la = smem + buf * S::BUF;
        tile_compute<AR, S, BK>(acc, la, la + AR::NPLANE * S::A_PLANE, wm, wn, lane);
        if (kt + 1 < nk) store_step(kt + 1, buf ^ 1);
        __syncthreads();
    }

    // beta == 0 decided once, outside the fragment loops: a conditional
    // load of old C inside them would put a vmcnt wait (which also waits
    // for the earlier stores) in front of every store
    Out* C = static_cast<Out*>(j.C);
    const int row0 = j.i0 + wm * S::WTM + 4 * (lane >> 4), col0 = j.j0 + wn * S::WTN + (lane & 15);
    if (j.beta0 != 0) {
#pragma unroll
        for (int fm = 0; fm < S::FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < S::FN; ++fn)
                gstore_frag<AR>(C, j.ldc, row0 + fm * 16, col0 + fn * 16, acc[fm][fn], j.s, true, j.m, j.n);
    } else {
#pragma unroll
        for (int fm = 0; fm < S::FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < S::FN; ++fn)
                gstore_frag<AR>(C, j.ldc, row0 + fm * 16, col0 + fn * 16, acc[fm][fn], j.s, false, j.m, j.n);
    }
}

template <class AR, int CFG>
struct GShape {
    static constexpr int BM = CFG == GCFG_32 ? 32 : CFG == GCFG_64 ? 64 : 128;
    static constexpr int BN = BM;
    static constexpr int BK = grouped_bk(AR::ID, CFG);
    static constexpr int LDS = TileShape<AR, true, true, BM, BN, BK, 2, 2>::LDS_BYTES;
};

template <class AR>
constexpr int g_lds_max()
{
    constexpr int a = GShape<AR, GCFG_32>::LDS, b = GShape<AR, GCFG_64>::LDS, c = GShape<AR, GCFG_128>::LDS;
    constexpr int ab = a > b ? a : b;
    return grouped_max_cfg(AR::ID) == GCFG_128 ? (ab > c ? ab : c) : ab;
}

template <class AR>
__device__ __forceinline__ void g_run(int cfg, const TileJob& j, uint32_t abytes, uint32_t bbytes, char* smem)
{
    if (cfg == GCFG_32) {
        using G = GShape<AR, GCFG_32>;
        gtile_job<AR, G::BM, G::BN, G::BK>(j, abytes, bbytes, smem);
        return;
    }
    if constexpr (grouped_max_cfg(AR::ID) == GCFG_128) {
        if (cfg == GCFG_128) {
            using G = GShape<AR, GCFG_128>;
            gtile_job<AR, G::BM, G::BN, G::BK>(j, abytes, bbytes, smem);
            return;
        }
    }
    using G = GShape<AR, GCFG_64>;
    gtile_job<AR, G::BM, G::BN, G::BK>(j, abytes, bbytes, smem);
}

// One wave's 32x32 block of a grouped instance (W32 items), operand layouts
// chosen at run time.
template <class AR>
__device__ __forceinline__ void g_run_wave(const TileJob& j)
{
    if (j.akc) {
        if (j.bkc) direct_job<AR, true, true, 32, 32>(j);
        else direct_job<AR, true, false, 32, 32>(j);
    } else {
        if (j.bkc) direct_job<AR, false, true, 32, 32>(j);
        else direct_job<AR, false, false, 32, 32>(j);
    }
}

// The plan's tables and the caller's pointer arrays do not change during a
// launch: read through the constant address space, at wave-uniform indices
// they become scalar loads, which wait on their own counter instead of
// behind the previous item's vector stores.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* g_const(const T* p)
{
    return (const __attribute__((address_space(4))) T*)(uintptr_t)p;
}
// Element idx of a table, word by word through the constant address space.
template <class T>
__device__ __forceinline__ T g_load(const T* base, long idx)
{
    static_assert(sizeof(T) % 4 == 0, "whole words");
    const __attribute__((address_space(4))) uint32_t* w = g_const(reinterpret_cast<const uint32_t*>(base + idx));
    uint32_t tmp[sizeof(T) / 4];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) tmp[i] = w[i];
    T out;
    __builtin_memcpy(&out, tmp, sizeof(T));
    return out;
}

__device__ __forceinline__ TileJob g_job(const GroupedDevice& d, const GItem& item, const GInst& in, const GGroup& g)
{
    TileJob j;
    j.A = static_cast<const char*>(g_const(d.A)[item.inst]);
    j.B = static_cast<const char*>(g_const(d.B)[item.inst]);
    j.C = g_const(d.C)[item.inst];
    j.lda = in.lda;
    j.ldb = in.ldb;
    j.ldc = in.ldc;
    j.m = in.m;
    j.n = in.n;
    j.k = in.k;
    j.i0 = item.i0;
    j.j0 = item.j0;
    j.s = Scalars{g.s.af, g.s.bf, g.s.ad, g.s.bd, g.s.ai, g.s.bi};
    j.beta0 = g.beta0;
    j.akc = g.akc;
    j.bkc = g.bkc;
    return j;
}

// FA: the family's first form; FB: its second (== FA for one-form families).
// Workgroup items first (the longer ones): workgroup b, on XCD b % 8, takes
// the items of queue q = b % GQ in turn with the other workgroups of that
// queue (rank r = b / GQ of nq: items r, r + nq, r + 2 nq, ... -- the queue
// is sorted longest first, so every workgroup gets one item of each stretch
// of similar cost).  Then wave items, the same way over all the grid's
// waves.  The assignment is fixed: no claims, so nothing waits on an atomic
// (whose return, on gfx9, retires in order behind the previous tile's
// stores) or on the barriers that shared a claimed number.
template <class FA, class FB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) grouped_gemm(const GroupedDevice d)
{
    constexpr int LA = g_lds_max<FA>(), LB = g_lds_max<FB>();
    __shared__ __attribute__((aligned(16))) char smem[LA > LB ? LA : LB];
    const int tid = threadIdx.x;
    // with fewer workgroups than queues, workgroup b also takes queues
    // b + G, b + 2 G, ... (G = gridDim.x) on its own
    const int G = (int)gridDim.x;
    const int qstep = G < GQ ? G : GQ;
    for (int q = (int)blockIdx.x % GQ; q < GQ; q += qstep) {
        const int nq = G < GQ ? 1 : (G - q + GQ - 1) / GQ;  // workgroups taking queue q
        const int rank = G < GQ ? 0 : (int)blockIdx.x / GQ;
        for (int it = d.qoff[q] + rank; it < d.qoff[q + 1]; it += nq) {
            const GItem item = g_load(d.items, it);
            const GInst in = g_load(d.insts, item.inst);
            const GGroup g = g_load(d.groups, in.group);
            const TileJob j = g_job(d, item, in, g);
            if (g.form == FA::ID) g_run<FA>(item.cfg, j, in.abytes, in.bbytes, smem);
            else g_run<FB>(item.cfg, j, in.abytes, in.bbytes, smem);
        }
    }
    if (d.nwitems > 0) {
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int nw = 4 * (int)gridDim.x;
        for (int w = 4 * (int)blockIdx.x + wave; w < d.nwitems; w += nw) {
            const GItem item = g_load(d.items, d.nitems + w);
            const GInst in = g_load(d.insts, item.inst);
            const GGroup g = g_load(d.groups, in.group);
            const TileJob j = g_job(d, item, in, g);
            if (g.form == FA::ID) g_run_wave<FA>(j);
            else g_run_wave<FB>(j);
        }
    }
}

template <class FA, class FB>
hipError_t run_grouped(const GroupedDevice& d, hipStream_t st)
{
    if (d.nitems + d.nwitems <= 0) return hipSuccess;
    auto kern = grouped_gemm<FA, FB>;
    const long slots = (long)kernel_occupancy(kern, 256) * device_cu_count();
    const long want = (long)d.nitems + ((long)d.nwitems + 3) / 4;  // four waves take wave items at once
    const long grid = want < slots ? want : slots;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, d);
    return hipGetLastError();
}

}  // namespace ssb
