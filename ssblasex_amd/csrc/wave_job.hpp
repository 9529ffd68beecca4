// wave_job.hpp -- one wave's block of C for the int8 and fp32 forms with
// operands loaded straight into MFMA fragment registers through buffer
// descriptors over the stored matrices (range-checked: a dword past the end
// of a matrix reads as zero, any leading dimension and alignment is read as
// whole 16-byte words), k-steps staged two register sets deep; and the
// fragment epilogues at any ldc and alignment.  Used by the grouped kernel's
// one-wave items and by the batched one-wave configs (direct16, direct32).
#pragma once
#include "gemm_kernels.hpp"

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


// A wave's FM x FN fragments stored column block by column block: the two
// 16-row halves of a column's 128-byte line (fragments 2q and 2q + 1 of one
// column block) leave back to back, so the L2 writes the line once.  beta == 0
// decided once, outside the loops: a conditional load of old C inside them
// would put a vmcnt wait (which also waits for the earlier stores) in front of
// every store.
template <class AR, int FM, int FN>
__device__ __forceinline__ void gstore_block(typename AR::Out* C, long ldc, int row0, int col0,
                                             const typename AR::Acc (&acc)[FM][FN], const Scalars& sc, bool beta0,
                                             int m, int n)
{
    if (beta0) {
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
                gstore_frag<AR>(C, ldc, row0 + fm * 16, col0 + fn * 16, acc[fm][fn], sc, true, m, n);
    } else {
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
                gstore_frag<AR>(C, ldc, row0 + fm * 16, col0 + fn * 16, acc[fm][fn], sc, false, m, n);
    }
}

// The old C values of one fragment (beta != 0) at any ldc and alignment,
// zero outside the matrix.
template <class AR>
__device__ __forceinline__ void gload_old_frag(const typename AR::Out* C, long ldc, int r0, int c, int m, int n,
                                               typename AR::Out (&old)[4])
{
    using Out = typename AR::Out;
    using u32x4_e = typename ElemAligned16<sizeof(Out)>::T;
    constexpr int NV = 4 * (int)sizeof(Out) / 16;
#pragma unroll
    for (int v = 0; v < 4; ++v) old[v] = Out(0);
    if (c >= n) return;
    const Out* src = C + (long)c * ldc + r0;
    if (r0 + 4 <= m) {
        u32x4_e w[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) w[i] = *gbl(reinterpret_cast<const u32x4_e*>(src) + i);
        __builtin_memcpy(old, w, sizeof(old));
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (r0 + v < m) old[v] = *gbl(src + v);
    }
}

// Epilogue of one fragment whose old C values were read ahead (beta != 0).
template <class AR>
__device__ __forceinline__ void gstore_frag_old(typename AR::Out* C, long ldc, int r0, int c,
                                                const typename AR::Acc& acc, const Scalars& sc, int m, int n,
                                                const typename AR::Out (&old)[4])
{
    using Out = typename AR::Out;
    using u32x4_e = typename ElemAligned16<sizeof(Out)>::T;
    constexpr int NV = 4 * (int)sizeof(Out) / 16;
    if (c >= n) return;
    Out* dst = C + (long)c * ldc + r0;
    Out res[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) res[v] = epi(acc[v], old[v], sc, false);
    if (r0 + 4 <= m) {
        u32x4_e w[NV];
        __builtin_memcpy(w, res, sizeof(res));
#pragma unroll
        for (int i = 0; i < NV; ++i) *gbl(reinterpret_cast<u32x4_e*>(dst) + i) = w[i];
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (r0 + v < m) *gbl(dst + v) = res[v];
    }
}

typedef int32_t gi32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) gi32x2 g_lds_i32x2;

// One operand of a wave tile (16 NF rows of op(X): NF 16-row MFMA fragments)
// read straight into fragment registers through a buffer descriptor over the
// stored matrix.  Lane l supplies row l%16 of a fragment and, within a k-step
// of KSTEP elements, the KL = KSTEP/4 elements from k = (l/16) KL on:
//  * K-contiguous: one 16-byte word per fragment; in the step that reaches
//    past K its elements at k >= K (the next row) are cleared;
//  * fp32 M/N-contiguous: four dwords per fragment, one per k; a dword at
//    k >= K starts past the end of the matrix and reads as zero;
//  * int8 M/N-contiguous: lane l reads the 16 bytes of rows 0..15 at k = l,
//    writes them to 1 KiB of the wave's LDS slot per fragment and reads its
//    fragment back with two ds_read_b64_tr_b8 (output lane o of a 16-lane
//    group receives column o of the 8 consecutive rows its group's lanes
//    point at).
// int8 words that cross the end of the matrix read as zero from the dword on
// that holds the last bytes; they are read again as the 16 bytes ending at
// the end and shifted (only where a lane of the wave holds one).
template <class AR, bool KC, int NF>
struct WaveOp {
    static_assert(NF == 1 || NF == 2, "one or two fragments");
    static constexpr int E = AR::E, KL = AR::KSTEP / 4;
    static_assert(KL * E == 16, "one 16-byte word per lane and fragment");

    // Byte offset of the word lane `lane` reads for fragment f at step kb.
    __device__ __forceinline__ static uint32_t off(int r0, int f, int kb, int ld, int lane, int s = 0)
    {
        if constexpr (KC) return (uint32_t)(((r0 + 16 * f + (lane & 15)) * ld + kb + (lane >> 4) * KL) * E);
        else if constexpr (E == 4) return (uint32_t)(((kb + (lane >> 4) * KL + s) * ld + r0 + 16 * f + (lane & 15)) * E);
        else return (uint32_t)((kb + lane) * ld + r0 + 16 * f);
    }

    __device__ __forceinline__ static void load(u32x4* w, __amdgpu_buffer_rsrc_t rs, int r0, int kb, int ld,
                                                int lane)
    {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            if constexpr (!KC && E == 4) {
#pragma unroll
                for (int s = 0; s < 4; ++s) w[f][s] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off(r0, f, kb, ld, lane, s), 0, 0);
            } else {
                w[f] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off(r0, f, kb, ld, lane), 0, 0);
            }
        }
    }

    // int8 words crossing the end of the matrix (`ext` bytes), and the k tail
    // of a K-contiguous operand (`kv` valid elements from kb on).
    __device__ __forceinline__ static void fix(u32x4* w, __amdgpu_buffer_rsrc_t rs, int r0, int kb, int ld, int lane,
                                               uint32_t ext, int kv)
    {
        if constexpr (E == 1) {
            bool any = false;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const uint32_t o = off(r0, f, kb, ld, lane);
                any |= o < ext && o + 16u > ext;
            }
            if (__any(any)) {
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const uint32_t o = off(r0, f, kb, ld, lane);
                    uint32_t d[4] = {w[f][0], w[f][1], w[f][2], w[f][3]};
                    if (o < ext && o + 16u > ext) {
                        if (ext < 16u) {
#pragma unroll
                            for (int i = 0; i < 4; ++i) d[i] = 0u;
                            for (uint32_t p = o; p < ext; ++p)
                                d[(p - o) >> 2] |= __builtin_amdgcn_raw_buffer_load_b8(rs, (int)p, 0, 0)
                                                   << (8 * ((p - o) & 3));
                        } else {
                            const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(ext - 16u), 0, 0);
                            const uint32_t sh = o - (ext - 16u), q = sh >> 2, r = sh & 3;
                            const uint32_t src[4] = {v[0], v[1], v[2], v[3]};
                            uint32_t t[5];
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                const uint32_t a = i + q;
                                t[i] = a == 0 ? src[0] : a == 1 ? src[1] : a == 2 ? src[2] : a == 3 ? src[3] : 0u;
                            }
                            t[4] = 0u;
#pragma unroll
                            for (int i = 0; i < 4; ++i) d[i] = __builtin_amdgcn_alignbyte(t[i + 1], t[i], r);
                        }
                    }
                    w[f] = u32x4{d[0], d[1], d[2], d[3]};
                }
            }
        }
        if constexpr (KC) {
            if (kv < AR::KSTEP) {
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    uint32_t d[4] = {w[f][0], w[f][1], w[f][2], w[f][3]};
                    clear_tail<16, E>(d, min(max(kv - (lane >> 4) * KL, 0), KL));
                    w[f] = u32x4{d[0], d[1], d[2], d[3]};
                }
            }
        }
    }

    // Words -> MFMA fragments (int8 M/N-contiguous: through the wave's 2 KiB LDS slot).
    __device__ __forceinline__ static void frags(u32x4* w, char* slot, int lane)
    {
        if constexpr (!KC && E == 1) {
#pragma unroll
            for (int f = 0; f < NF; ++f) *reinterpret_cast<u32x4*>(slot + 1024 * f + 16 * lane) = w[f];
            const int r = lane & 15, g = lane >> 4;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                gi32x2 v[2];
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    v[h] = __builtin_amdgcn_ds_read_tr8_b64_v2i32(
                        (g_lds_i32x2*)(slot + 1024 * f + 16 * (16 * g + 8 * h + (r >> 1)) + 8 * (r & 1)));
                w[f] = u32x4{(uint32_t)v[0].x, (uint32_t)v[0].y, (uint32_t)v[1].x, (uint32_t)v[1].y};
            }
        }
    }
};

// One wave's (16 FM) x (16 FN) block of C for the int8 and fp32 forms:
// fragments loaded straight from global memory through buffer descriptors
// (abytes / bbytes: extent of the stored A / B), NS k-steps per stage, the
// next stage's loads in flight while the current one is multiplied (two
// register sets, the loop unrolled by two so no set is copied while its
// loads are in flight).  PREFETCH_C: for beta != 0 the old C values are
// read before the k-loop instead of in the epilogue.  `slot`: 4 KiB of LDS
// private to the wave (int8 M/N-contiguous operands pass through it).
template <class AR, bool AKC, bool BKC, int FM, int FN, int NS, bool PREFETCH_C>
__device__ __forceinline__ void wave_job(const TileJob& j, uint32_t abytes, uint32_t bbytes, char* slot)
{
    using Out = typename AR::Out;
    using OA = WaveOp<AR, AKC, FM>;
    using OB = WaveOp<AR, BKC, FN>;
    constexpr int KS = AR::KSTEP;
    int lane = threadIdx.x & 63;
    __asm__ volatile("" : "+v"(lane));
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(j.A, abytes);
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(j.B, bbytes);
    const int lda = (int)j.lda, ldb = (int)j.ldb;
    const int nsteps = (j.k + KS - 1) / KS;
    const int nst = (nsteps + NS - 1) / NS;
    Out* C = static_cast<Out*>(j.C);
    const int row0 = j.i0 + 4 * (lane >> 4), col0 = j.j0 + (lane & 15);

    typename AR::Acc acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = typename AR::Acc{};
    Out old[PREFETCH_C ? FM : 1][PREFETCH_C ? FN : 1][4];
    if constexpr (PREFETCH_C) {
        if (j.beta0 == 0) {
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn)
                    gload_old_frag<AR>(C, j.ldc, row0 + fm * 16, col0 + fn * 16, j.m, j.n, old[fm][fn]);
        }
    }

    u32x4 a0[NS][FM], b0[NS][FN], a1[NS][FM], b1[NS][FN];
    auto load = [&](u32x4 (*a)[FM], u32x4 (*b)[FN], int s) {
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const int kb = (s * NS + t) * KS;
            if (kb < j.k) {
                OA::load(a[t], ra, j.i0, kb, lda, lane);
                OB::load(b[t], rb, j.j0, kb, ldb, lane);
            }
        }
    };
    auto run = [&](u32x4 (*a)[FM], u32x4 (*b)[FN], int s) {
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const int kb = (s * NS + t) * KS;
            if (kb < j.k) {
                OA::fix(a[t], ra, j.i0, kb, lda, lane, abytes, j.k - kb);
                OB::fix(b[t], rb, j.j0, kb, ldb, lane, bbytes, j.k - kb);
                OA::frags(a[t], slot, lane);
                OB::frags(b[t], slot + 2048, lane);
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn) AR::mma(acc[fm][fn], &a[t][fm], &b[t][fn]);
            }
        }
    };
    if (nst > 0) load(a0, b0, 0);
    for (int s = 0; s < nst; s += 2) {
        if (s + 1 < nst) load(a1, b1, s + 1);
        run(a0, b0, s);
        if (s + 1 >= nst) break;
        if (s + 2 < nst) load(a0, b0, s + 2);
        run(a1, b1, s + 1);
    }

    if constexpr (PREFETCH_C) {
        if (j.beta0 == 0) {
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
                    gstore_frag_old<AR>(C, j.ldc, row0 + fm * 16, col0 + fn * 16, acc[fm][fn], j.s, j.m, j.n,
                                        old[fm][fn]);
            return;
        }
    }
    gstore_block<AR, FM, FN>(C, j.ldc, row0, col0, acc, j.s, j.beta0 != 0, j.m, j.n);
}

// Bytes a buffer descriptor over a stored matrix spans (rows x cols at
// leading dimension ld), 0 when it -- with the reach of a tile past its
// edge -- does not fit 31-bit offsets.
__host__ __device__ __forceinline__ uint32_t wave_extent(long rows, long cols, long ld, int e)
{
    const long ext = ((cols - 1) * ld + rows) * (long)e;
    const long reach = ext + 256L * (ld + 1) * e;
    return reach >= (1L << 31) ? 0u : (uint32_t)ext;
}

// Batched kernel of one-wave WTM x WTN blocks for the int8 and fp32 forms
// (configs direct16 / direct32): waves walk the (instance, block) work ids.
// abytes / bbytes: the extent of every instance's stored A / B (the same for
// all instances of a batch).
template <class AR, bool AKC, bool BKC, int WTM, int WTN, int NS>
__global__ void __launch_bounds__(256) wave_gemm(const KParams p, uint32_t abytes, uint32_t bbytes)
{
    __shared__ __attribute__((aligned(16))) char smem[4 * 4096];
    const int per_inst = p.tiles_m * p.tiles_n;
    const long nwaves = (long)gridDim.x * 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    char* slot = smem + 4096 * wave;
    for (long t = (long)blockIdx.x * 4 + wave; t < p.total; t += nwaves) {
        const int inst = (int)(t / per_inst);
        const int rem = (int)(t % per_inst);
        const TileJob j = make_job(p, inst, (rem % p.tiles_m) * WTM, (rem / p.tiles_m) * WTN);
        wave_job<AR, AKC, BKC, WTM / 16, WTN / 16, NS, true>(j, abytes, bbytes, slot);
    }
}

}  // namespace ssb
