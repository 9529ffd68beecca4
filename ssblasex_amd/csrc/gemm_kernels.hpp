// gemm_kernels.hpp -- the two kernel families of the batched GEMM.
//
// tile_gemm: LDS-tiled persistent kernel for instances of 64x64 and up.
//   One workgroup = WAVES_M x WAVES_N waves computes one BM x BN output
//   tile of one instance per iteration of its persistent loop over
//   (instance, tile) work ids.  A and B tiles are staged global -> registers
//   -> LDS with two LDS buffers (the next k-block's global loads are in
//   flight while the current block's MFMAs run).  The staging absorbs the
//   OP_N / OP_T layouts: a K-contiguous operand is copied 16 bytes at a time,
//   an M/N-contiguous operand is loaded as small 2-D blocks and transposed
//   in registers so that LDS always holds [row][k] with k contiguous.
//   Epilogue: C = alpha*acc + beta*C fused, C not read when beta == 0.
//
// direct_gemm: one wave per WTM x WTN block, operands loaded straight from
//   global memory into MFMA registers (no LDS, no barriers).  For small
//   instances (both m and n at most 64) where a tile would be mostly empty.
//
// Operand view: op(A) element (i, k) sits at A[i*sra + k*ska] and op(B)
// element (k, j) at B[j*srb + k*skb] (element strides).  OP_T for A and
// OP_N for B make k the contiguous index (ska == 1 / skb == 1).
#pragma once
#include "device_common.hpp"

namespace ssb {

struct KParams {
    const void* const* A;
    const void* const* B;
    void* const* C;
    long lda, ldb, ldc;
    int m, n, k;
    int batch;
    int tiles_m, tiles_n;  // per instance
    long total;            // batch * tiles_m * tiles_n
    Scalars s;
    int beta0;             // beta == 0: C is write-only
    int ld_vec_a, ld_vec_b, ld_vec_c;  // leading dimensions allow 16-byte vectors
};

// ---------------------------------------------------------------------------
// Element access helpers
// ---------------------------------------------------------------------------
template <int E>
struct Raw;
template <> struct Raw<1> { using T = uint8_t; };
template <> struct Raw<4> { using T = uint32_t; };
template <> struct Raw<8> { using T = uint64_t; };

template <int E>
__device__ __forceinline__ typename Raw<E>::T load_elem(const char* base, long off_elems)
{
    return *gbl(reinterpret_cast<const typename Raw<E>::T*>(base + off_elems * E));
}

// ---------------------------------------------------------------------------
// Stager: one operand tile (ROWS x BK input elements) global -> regs -> LDS.
//
// The tile is cut into units, one or more per thread.  A unit is read with
// "words" (one global load each: 16 bytes, or 4 bytes for an M/N-contiguous
// int8 operand) and written to LDS as [row][k] rows:
//   K-contiguous operand:  unit = one word = WE consecutive k of one row.
//   M/N-contiguous:        unit = RU rows x KU k, one word per k (the RU
//                          rows are contiguous in memory); transposed in
//                          registers before the LDS store.
// ---------------------------------------------------------------------------
template <class AR, bool KC, int ROWS, int BK, int NT>
struct StagerImpl {
    static constexpr int E = AR::E;
    static constexpr int WB = (!KC && E == 1) ? 4 : 16;  // bytes per word
    static constexpr int WE = WB / E;                     // elements per word
    static constexpr int WD = WB / 4;                     // dwords per word
    static constexpr int RU = KC ? 1 : WE;                // rows per unit
    static constexpr int KU = KC ? WE : (E == 1 ? 8 : WE); // k per unit
    static constexpr int NW = KC ? 1 : KU;                // words per unit
    static constexpr int RG = ROWS / RU;
    static constexpr int KG = BK / KU;
    static constexpr int UNITS = RG * KG;
    static constexpr int UPT = (UNITS + NT - 1) / NT;
    static constexpr int UD = NW * WD;                    // dwords per unit
    static constexpr int RAWD = UPT * UD;                 // dwords per thread
    static_assert(BK % KU == 0, "BK must hold whole units");
    static_assert(ROWS % RU == 0, "rows must hold whole units");

    __device__ __forceinline__ static void unit_coords(int u, int& r0, int& k0)
    {
        if constexpr (KC) { k0 = (u % KG) * KU; r0 = u / KG; }
        else { r0 = (u % RG) * RU; k0 = (u / RG) * KU; }
    }
    __device__ __forceinline__ static bool active(int u) { return (UNITS % NT) == 0 || u < UNITS; }

    // address (bytes from base) of word w of a unit
    __device__ __forceinline__ static long word_off(int r0, int k0, int w, long sr, long sk)
    {
        if constexpr (KC) return ((long)r0 * sr + k0) * E;
        else return ((long)r0 + (long)(k0 + w) * sk) * E;
    }

    __device__ __forceinline__ static void put_word(uint32_t* raw, int x, int w, const char* src)
    {
        if constexpr (WB == 16) {
            u32x4 v = *gbl(reinterpret_cast<const u32x4*>(src));
            raw[x * UD + w * 4 + 0] = v.x;
            raw[x * UD + w * 4 + 1] = v.y;
            raw[x * UD + w * 4 + 2] = v.z;
            raw[x * UD + w * 4 + 3] = v.w;
        } else {
            raw[x * UD + w] = *gbl(reinterpret_cast<const uint32_t*>(src));
        }
    }

    // Every row and k of this tile step in range, words aligned.
    __device__ __forceinline__ static void load_fast(uint32_t* raw, const char* base, long sr, long sk, int tid)
    {
#pragma unroll
        for (int x = 0; x < UPT; ++x) {
            const int u = tid + x * NT;
            if (!active(u)) break;
            int r0, k0;
            unit_coords(u, r0, k0);
#pragma unroll
            for (int w = 0; w < NW; ++w) put_word(raw, x, w, base + word_off(r0, k0, w, sr, sk));
        }
    }

    // Zero outside rows_valid x k_valid; a word is read with one vector load
    // when it is whole and `vec`, else element by element.
    __device__ __forceinline__ static void load_guarded(uint32_t* raw, const char* base, long sr, long sk,
                                                        int rows_valid, int k_valid, bool vec, int tid)
    {
        using RT = typename Raw<E>::T;
#pragma unroll
        for (int x = 0; x < UPT; ++x) {
            const int u = tid + x * NT;
            if (!active(u)) break;
            int r0, k0;
            unit_coords(u, r0, k0);
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                bool whole;
                if constexpr (KC) whole = r0 < rows_valid && k0 + WE <= k_valid;
                else whole = r0 + RU <= rows_valid && k0 + w < k_valid;
                if (vec && whole) {
                    put_word(raw, x, w, base + word_off(r0, k0, w, sr, sk));
                } else {
                    RT vals[WE];
#pragma unroll
                    for (int e = 0; e < WE; ++e) {
                        bool ok;
                        long off;
                        if constexpr (KC) {
                            ok = r0 < rows_valid && k0 + e < k_valid;
                            off = (long)r0 * sr + (k0 + e);
                        } else {
                            ok = r0 + e < rows_valid && k0 + w < k_valid;
                            off = (long)(r0 + e) + (long)(k0 + w) * sk;
                        }
                        vals[e] = ok ? load_elem<E>(base, off) : RT(0);
                    }
                    uint32_t d[WD];
                    __builtin_memcpy(d, vals, WB);
#pragma unroll
                    for (int i = 0; i < WD; ++i) raw[x * UD + w * WD + i] = d[i];
                }
            }
        }
    }

    // Write the staged units to LDS planes ([row][k], pitch rowb bytes).
    __device__ __forceinline__ static void store(const uint32_t* raw, char* lds, int plane_bytes, int rowb, int tid)
    {
#pragma unroll
        for (int x = 0; x < UPT; ++x) {
            const int u = tid + x * NT;
            if (!active(u)) break;
            int r0, k0;
            unit_coords(u, r0, k0);
            const uint32_t* r = raw + x * UD;
            // rows[i]: the KU*E bytes of row r0+i (16 bytes, or 8 for int8 M/N-contiguous)
            u32x4 rows[RU];
            if constexpr (KC) {
                rows[0] = u32x4{r[0], r[1], r[2], r[3]};
            } else if constexpr (E == 1) {
                uint32_t o0[4], o1[4];
                transpose4x4_bytes(&r[0], o0);  // k0..k0+3
                transpose4x4_bytes(&r[4], o1);  // k0+4..k0+7
#pragma unroll
                for (int i = 0; i < 4; ++i) rows[i] = u32x4{o0[i], o1[i], 0u, 0u};
            } else if constexpr (E == 4) {
#pragma unroll
                for (int i = 0; i < 4; ++i) rows[i] = u32x4{r[i], r[4 + i], r[8 + i], r[12 + i]};
            } else {  // E == 8: word w = rows r0, r0+1 at k0+w
                rows[0] = u32x4{r[0], r[1], r[4], r[5]};
                rows[1] = u32x4{r[2], r[3], r[6], r[7]};
            }
            const int kbyte = k0 * AR::LDS_E;
#pragma unroll
            for (int i = 0; i < RU; ++i) {
                char* dst = lds + lds_row_off<AR>(r0 + i, kbyte, rowb);
                if constexpr (!KC && E == 1) {
                    *reinterpret_cast<u32x2*>(dst) = u32x2{rows[i][0], rows[i][1]};
                } else if constexpr (AR::LDS_E == E) {
                    *reinterpret_cast<u32x4*>(dst) = rows[i];
                } else {
                    u32x4 planes[AR::NPLANE];
                    AR::to_planes(rows[i], planes);
#pragma unroll
                    for (int pl = 0; pl < AR::NPLANE; ++pl)
                        *reinterpret_cast<u32x2*>(dst + pl * plane_bytes) = u32x2{planes[pl][0], planes[pl][1]};
                }
            }
        }
    }
};

// Staging registers of one operand.  KCM: 1 = K-contiguous, 0 = M/N-
// contiguous (fixed at compile time), 2 = chosen per call by the `kc`
// argument (one register set sized for either layout; used where one
// kernel serves instances of both layouts).
template <class AR, int KCM, int ROWS, int BK, int NT>
struct Stager {
    using IK = StagerImpl<AR, true, ROWS, BK, NT>;
    using IN = StagerImpl<AR, false, ROWS, BK, NT>;
    static constexpr int RAWD = KCM == 1 ? IK::RAWD : KCM == 0 ? IN::RAWD : (IK::RAWD > IN::RAWD ? IK::RAWD : IN::RAWD);
    uint32_t raw[RAWD];

    __device__ __forceinline__ static bool kc_of(bool kc) { return KCM == 2 ? kc : KCM == 1; }
    // bytes per global load word for the layout
    __device__ __forceinline__ static int word_bytes(bool kc) { return kc_of(kc) ? IK::WB : IN::WB; }

    __device__ __forceinline__ void load_fast(bool kc, const char* base, long sr, long sk, int tid)
    {
        if constexpr (KCM == 1) IK::load_fast(raw, base, sr, sk, tid);
        else if constexpr (KCM == 0) IN::load_fast(raw, base, sr, sk, tid);
        else if (kc) IK::load_fast(raw, base, sr, sk, tid);
        else IN::load_fast(raw, base, sr, sk, tid);
    }
    __device__ __forceinline__ void load_guarded(bool kc, const char* base, long sr, long sk, int rows_valid,
                                                 int k_valid, bool vec, int tid)
    {
        if constexpr (KCM == 1) IK::load_guarded(raw, base, sr, sk, rows_valid, k_valid, vec, tid);
        else if constexpr (KCM == 0) IN::load_guarded(raw, base, sr, sk, rows_valid, k_valid, vec, tid);
        else if (kc) IK::load_guarded(raw, base, sr, sk, rows_valid, k_valid, vec, tid);
        else IN::load_guarded(raw, base, sr, sk, rows_valid, k_valid, vec, tid);
    }
    __device__ __forceinline__ void store(bool kc, char* lds, int plane_bytes, int rowb, int tid) const
    {
        if constexpr (KCM == 1) IK::store(raw, lds, plane_bytes, rowb, tid);
        else if constexpr (KCM == 0) IN::store(raw, lds, plane_bytes, rowb, tid);
        else if (kc) IK::store(raw, lds, plane_bytes, rowb, tid);
        else IN::store(raw, lds, plane_bytes, rowb, tid);
    }
};

// ---------------------------------------------------------------------------
// Epilogue for one 16x16 accumulator fragment.
//   r0: first of the lane's 4 consecutive rows, c: the lane's column.
// ---------------------------------------------------------------------------
template <class AR>
__device__ __forceinline__ void store_frag(typename AR::Out* C, long ldc, int r0, int c,
                                           const typename AR::Acc& acc, const Scalars& sc,
                                           bool beta0, bool fast, int m, int n)
{
    using Out = typename AR::Out;
    if (fast) {
        Out* dst = C + (long)c * ldc + r0;
        if constexpr (sizeof(Out) == 4) {
            // Elements go through plain arrays and memcpy: hipcc 7.2 folds
            // __builtin_bit_cast of ext-vector elements in this unrolled
            // loop into one reused dword.
            Out old[4] = {Out(0), Out(0), Out(0), Out(0)};
            if (!beta0) {
                u32x4 t = *gbl(reinterpret_cast<const u32x4*>(dst));
                uint32_t w[4] = {t.x, t.y, t.z, t.w};
                __builtin_memcpy(old, w, 16);
            }
            Out res[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) res[v] = epi(acc[v], old[v], sc, beta0);
            uint32_t w[4];
            __builtin_memcpy(w, res, 16);
            *gbl(reinterpret_cast<u32x4*>(dst)) = u32x4{w[0], w[1], w[2], w[3]};
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                f64x2 old = {0.0, 0.0};
                if (!beta0) old = *gbl(reinterpret_cast<const f64x2*>(dst + 2 * h));
                f64x2 res;
                res[0] = epi((double)acc[2 * h], old[0], sc, beta0);
                res[1] = epi((double)acc[2 * h + 1], old[1], sc, beta0);
                *gbl(reinterpret_cast<f64x2*>(dst + 2 * h)) = res;
            }
        }
    } else {
        if (c >= n) return;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int r = r0 + v;
            if (r < m) {
                Out* dst = C + (long)c * ldc + r;
                Out o = beta0 ? Out(0) : *gbl(dst);
                *gbl(dst) = epi(acc[v], o, sc, beta0);
            }
        }
    }
}

// 16 bytes (or two times 16) of results / old values of one fragment lane:
// the lane's 4 consecutive rows of one column, 16-byte aligned.
template <class Out>
__device__ __forceinline__ void store_vec4(Out* dst, const Out (&res)[4])
{
    if constexpr (sizeof(Out) == 4) {
        uint32_t w[4];
        __builtin_memcpy(w, res, 16);
        *gbl(reinterpret_cast<u32x4*>(dst)) = u32x4{w[0], w[1], w[2], w[3]};
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) *gbl(reinterpret_cast<f64x2*>(dst + 2 * h)) = f64x2{res[2 * h], res[2 * h + 1]};
    }
}
template <class Out>
__device__ __forceinline__ void load_vec4(const Out* src, Out (&old)[4])
{
    if constexpr (sizeof(Out) == 4) {
        const u32x4 t = *gbl(reinterpret_cast<const u32x4*>(src));
        const uint32_t w[4] = {t.x, t.y, t.z, t.w};
        __builtin_memcpy(old, w, 16);
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const f64x2 t = *gbl(reinterpret_cast<const f64x2*>(src + 2 * h));
            old[2 * h] = t[0];
            old[2 * h + 1] = t[1];
        }
    }
}

// Fused alpha/beta epilogue of a wave's FM x FN fragments; pos(fm, fn, r0, c)
// gives the lane's first row and its column in fragment (fm, fn).  The
// beta == 0 / beta != 0 and whole / edge choices are made once for the wave,
// outside the fragment loops: a load issued conditionally inside the loop
// puts a vmcnt wait in front of every store, and on gfx9 that wait also
// waits for every store issued before it, so the tile's stores would leave
// one round trip at a time.  With beta != 0 the old values are read D
// fragments ahead through a ring of D fragments: fragment f + D is requested
// once f's results are computed and before they are stored, so the wait for
// an old value never covers a store.  fp64 results keep the per-fragment
// routine: the split epilogue made the fp64 128x128 tile (at the edge of
// its register budget) 20 % slower, for no gain in its stores.
// Returns the number of vector memory instructions the wave issued (F
// stores for beta == 0, F loads and F stores for beta != 0, F = FM * FN) on
// the two whole-tile paths, -1 on the others (count not fixed).
// WT (4-byte results) and `through`: the whole-tile stores are buffer
// stores with the sc0 sc1 policy (system scope: written through L2), so a
// kernel that ends with these stores leaves no dirty C lines for the
// end-of-kernel L2 write-back to flush (int8 512^3 x 64, one 256x256 tile
// per workgroup: 20.7 instead of 23.3 µs per launch).  For the last tile of
// a workgroup only: a tile followed by more work keeps write-back stores
// (through them, 512^3 x 256 went from 95.7 to 105.9 µs).  Used when the
// instance's C spans less than 2 GiB (32-bit byte offsets).
template <class AR, int FM, int FN, bool WT = false, class Pos>
__device__ __forceinline__ int store_tile(typename AR::Out* C, long ldc, const typename AR::Acc (&acc)[FM][FN],
                                           const Scalars& sc, bool beta0, bool fast, int m, int n, Pos pos,
                                           bool through = false)
{
    using Out = typename AR::Out;
    constexpr int F = FM * FN;
    constexpr bool WTB = WT && sizeof(Out) == 4;
    const long cbytes = ((long)(n - 1) * ldc + m) * (long)sizeof(Out);
    const bool wt = WTB && through && cbytes < 0x7fffffffL;
    __amdgpu_buffer_rsrc_t crs = make_rsrc(C, wt ? (uint32_t)cbytes : 0u);
    typedef int i32x4v __attribute__((ext_vector_type(4)));
    auto put = [&](int r0, int c, const Out(&res)[4]) {
        if constexpr (WTB) {
            if (wt) {
                uint32_t w[4];
                __builtin_memcpy(w, res, 16);
                __builtin_amdgcn_raw_buffer_store_b128(i32x4v{(int)w[0], (int)w[1], (int)w[2], (int)w[3]}, crs,
                                                       (int)(((long)c * ldc + r0) * (long)sizeof(Out)), 0, 17);
                return;
            }
        }
        store_vec4<Out>(C + (long)c * ldc + r0, res);
    };
    if constexpr (sizeof(Out) == 8) {
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) {
                int r0, c;
                pos(fm, fn, r0, c);
                store_frag<AR>(C, ldc, r0, c, acc[fm][fn], sc, beta0, fast, m, n);
            }
        return -1;
    }
    // Visit order: with FM even, fragment rows 2q and 2q + 1 of one column
    // block go out back to back, so the two 64-byte halves that a column's
    // lanes write of one 128-byte line leave together.
    auto fm_of = [](int f) { return FM % 2 ? f / FN : 2 * (f / (2 * FN)) + (f % 2); };
    auto fn_of = [](int f) { return FM % 2 ? f % FN : (f % (2 * FN)) / 2; };
    if (fast && beta0) {
#pragma unroll
        for (int f = 0; f < F; ++f) {
            int r0, c;
            pos(fm_of(f), fn_of(f), r0, c);
            Out res[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) res[v] = epi((Out)acc[fm_of(f)][fn_of(f)][v], Out(0), sc, true);
            put(r0, c, res);
        }
        return F;
    } else if (fast) {
        constexpr int D = F < 4 ? F : 4;
        Out old[D][4];
        auto load = [&](int f) {
            int r0, c;
            pos(fm_of(f), fn_of(f), r0, c);
            load_vec4<Out>(C + (long)c * ldc + r0, old[f % D]);
        };
#pragma unroll
        for (int f = 0; f < D; ++f) load(f);
#pragma unroll
        for (int f = 0; f < F; ++f) {
            int r0, c;
            pos(fm_of(f), fn_of(f), r0, c);
            Out res[4];
#pragma unroll
            for (int v = 0; v < 4; ++v)
                res[v] = epi((Out)acc[fm_of(f)][fn_of(f)][v], old[f % D][v], sc, false);
            if (f + D < F) load(f + D);
            put(r0, c, res);
        }
        return 2 * F;
    } else {
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) {
                int r0, c;
                pos(fm, fn, r0, c);
                store_frag<AR>(C, ldc, r0, c, acc[fm][fn], sc, beta0, false, m, n);
            }
        return -1;
    }
}

// The old C values of one fragment (the lane's 4 rows of column c), read
// before the k-loop of a small tile when beta != 0, so that their latency
// overlaps the operand loads instead of following the last MFMA.  Zero
// outside the matrix.
template <class AR>
__device__ __forceinline__ void load_old_frag(const typename AR::Out* C, long ldc, int r0, int c, bool fast, int m,
                                              int n, typename AR::Out (&old)[4])
{
    using Out = typename AR::Out;
    const Out* src = C + (long)c * ldc + r0;
    if (fast) {
        if constexpr (sizeof(Out) == 4) {
            const u32x4 t = *gbl(reinterpret_cast<const u32x4*>(src));
            const uint32_t w[4] = {t.x, t.y, t.z, t.w};
            __builtin_memcpy(old, w, 16);
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const f64x2 t = *gbl(reinterpret_cast<const f64x2*>(src + 2 * h));
                old[2 * h] = t[0];
                old[2 * h + 1] = t[1];
            }
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) old[v] = (c < n && r0 + v < m) ? *gbl(src + v) : Out(0);
    }
}

// Epilogue of one fragment whose old C values were read ahead (beta != 0).
template <class AR>
__device__ __forceinline__ void store_frag_old(typename AR::Out* C, long ldc, int r0, int c,
                                               const typename AR::Acc& acc, const Scalars& sc, bool fast, int m, int n,
                                               const typename AR::Out (&old)[4])
{
    using Out = typename AR::Out;
    Out* dst = C + (long)c * ldc + r0;
    Out res[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) res[v] = epi(acc[v], old[v], sc, false);
    if (fast) {
        if constexpr (sizeof(Out) == 4) {
            uint32_t w[4];
            __builtin_memcpy(w, res, 16);
            *gbl(reinterpret_cast<u32x4*>(dst)) = u32x4{w[0], w[1], w[2], w[3]};
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) *gbl(reinterpret_cast<f64x2*>(dst + 2 * h)) = f64x2{res[2 * h], res[2 * h + 1]};
        }
    } else {
        if (c >= n) return;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (r0 + v < m) *gbl(dst + v) = res[v];
    }
}

// ---------------------------------------------------------------------------
// One output tile of one instance: what a work item hands to a tile routine.
// ---------------------------------------------------------------------------
struct TileJob {
    const char* A;  // instance base pointers
    const char* B;
    void* C;
    long lda, ldb, ldc;
    int m, n, k;
    int i0, j0;     // tile origin in C
    Scalars s;
    int beta0;
    int akc, bkc;   // operand layouts, read when a tile routine is compiled for either
};

template <class AR, bool AKC, bool BKC, int BM, int BN, int BK, int WAVES_M, int WAVES_N>
struct TileShape {
    static constexpr int NT = 64 * WAVES_M * WAVES_N;
    static constexpr int ROWB = lds_row_pitch<AR>(BK);  // LDS row pitch
    static constexpr int A_PLANE = BM * ROWB;
    static constexpr int B_PLANE = BN * ROWB;
    static constexpr int BUF = AR::NPLANE * (A_PLANE + B_PLANE);
    static constexpr int LDS_BYTES = 2 * BUF;
    static constexpr int WTM = BM / WAVES_M, WTN = BN / WAVES_N;
    static constexpr int FM = WTM / 16, FN = WTN / 16;
    static_assert(WTM % 16 == 0 && WTN % 16 == 0, "wave tile must be whole 16x16 fragments");
    static_assert(BK % AR::KSTEP == 0, "BK must hold whole MFMA k-steps");
};

// MFMA work of one LDS k-block: the wave's FM x FN fragments over BK.
// la / lb: the A and B planes of the k-block ([row][k], pitch S::ROWB).
template <class AR, class S, int BK>
__device__ __forceinline__ void tile_compute(typename AR::Acc (&acc)[S::FM][S::FN], const char* la, const char* lb,
                                             int wm, int wn, int lane)
{
    constexpr int KC_L = 16 / AR::LDS_E;  // LDS elements per 16-byte fragment chunk
#pragma unroll
    for (int ks = 0; ks < BK / AR::KSTEP; ++ks) {
        const int kbyte = (ks * AR::KSTEP + (lane >> 4) * KC_L) * AR::LDS_E;
        u32x4 af[S::FM][AR::NPLANE], bfg[S::FN][AR::NPLANE];
#pragma unroll
        for (int fm = 0; fm < S::FM; ++fm) {
            const int row = wm * S::WTM + fm * 16 + frag_row<AR>(lane & 15);
#pragma unroll
            for (int pl = 0; pl < AR::NPLANE; ++pl)
                af[fm][pl] = *reinterpret_cast<const u32x4*>(la + pl * S::A_PLANE + lds_row_off<AR>(row, kbyte, S::ROWB));
        }
#pragma unroll
        for (int fn = 0; fn < S::FN; ++fn) {
            const int row = wn * S::WTN + fn * 16 + (lane & 15);
#pragma unroll
            for (int pl = 0; pl < AR::NPLANE; ++pl)
                bfg[fn][pl] = *reinterpret_cast<const u32x4*>(lb + pl * S::B_PLANE + lds_row_off<AR>(row, kbyte, S::ROWB));
        }
#pragma unroll
        for (int fm = 0; fm < S::FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < S::FN; ++fn) AR::mma(acc[fm][fn], af[fm], bfg[fn]);
    }
}

// LDS-staged workgroup tile: BM x BN of C, all threads of the workgroup.
// smem must hold TileShape::LDS_BYTES; the caller's previous use of it must
// be finished by every thread (the routine starts with a barrier).
//
// STAGES = 1: the global loads of k-block kt+1 are in flight while block kt
// is multiplied.  STAGES = 2: two register staging sets, loads issued two
// k-blocks ahead (covers twice the memory latency for twice the staging
// registers).
// AKM / BKM: operand layout mode, 1 = K-contiguous, 0 = M/N-contiguous,
// 2 = taken from j.akc / j.bkc at run time.
template <class AR, int AKM, int BKM, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int STAGES = 1>
__device__ __forceinline__ void tile_job(const TileJob& j, char* smem)
{
    using S = TileShape<AR, true, true, BM, BN, BK, WAVES_M, WAVES_N>;
    using Out = typename AR::Out;
    constexpr int NT = S::NT;
    constexpr int E = AR::E;
    using SA = Stager<AR, AKM, BM, BK, NT>;
    using SB = Stager<AR, BKM, BN, BK, NT>;
    const bool AKC = AKM == 2 ? (j.akc != 0) : (AKM == 1);
    const bool BKC = BKM == 2 ? (j.bkc != 0) : (BKM == 1);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    // element strides of the operand views
    const long sra = AKC ? j.lda : 1, ska = AKC ? 1 : j.lda;
    const long srb = BKC ? j.ldb : 1, skb = BKC ? 1 : j.ldb;
    const int nk = (j.k + BK - 1) / BK;
    const int rows_a = min(BM, j.m - j.i0);
    const int rows_b = min(BN, j.n - j.j0);
    const int a_wb = SA::word_bytes(AKC), b_wb = SB::word_bytes(BKC);
    const bool a_vec = ((j.lda * E) % a_wb) == 0 && (reinterpret_cast<uintptr_t>(j.A) % a_wb) == 0;
    const bool b_vec = ((j.ldb * E) % b_wb) == 0 && (reinterpret_cast<uintptr_t>(j.B) % b_wb) == 0;
    const bool a_fast = a_vec && rows_a == BM;
    const bool b_fast = b_vec && rows_b == BN;
    // tile origin (row i0 / j0, k 0) in bytes
    const char* Ab = j.A + (long)j.i0 * sra * E;
    const char* Bb = j.B + (long)j.j0 * srb * E;

    SA sa[STAGES];
    SB sb[STAGES];
    typename AR::Acc acc[S::FM][S::FN];
#pragma unroll
    for (int fm = 0; fm < S::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < S::FN; ++fn) acc[fm][fn] = typename AR::Acc{};

    // small wave tiles read the old C values (beta != 0) before the k-loop
    constexpr bool PRE = S::FM * S::FN <= 4;
    Out* C = static_cast<Out*>(j.C);
    const bool c_fast = rows_a == BM && rows_b == BN && ((j.ldc * (long)sizeof(Out)) % 16) == 0 && aligned16(C);
    Out old[PRE ? S::FM : 1][PRE ? S::FN : 1][4];
    if (PRE && j.beta0 == 0) {
#pragma unroll
        for (int fm = 0; fm < (PRE ? S::FM : 1); ++fm)
#pragma unroll
            for (int fn = 0; fn < (PRE ? S::FN : 1); ++fn)
                load_old_frag<AR>(C, j.ldc, j.i0 + wm * S::WTM + fm * 16 + 4 * (lane >> 4),
                                  j.j0 + wn * S::WTN + fn * 16 + (lane & 15), c_fast, j.m, j.n, old[fm][fn]);
    }

    auto load_step = [&](int kt, SA& xa, SB& xb) {
        const int k0 = kt * BK;
        const int kv = j.k - k0;
        const char* ak = Ab + (long)k0 * ska * E;
        const char* bk = Bb + (long)k0 * skb * E;
        if (a_fast && kv >= BK) xa.load_fast(AKC, ak, sra, ska, tid);
        else xa.load_guarded(AKC, ak, sra, ska, rows_a, kv, a_vec, tid);
        if (b_fast && kv >= BK) xb.load_fast(BKC, bk, srb, skb, tid);
        else xb.load_guarded(BKC, bk, srb, skb, rows_b, kv, b_vec, tid);
    };
    auto store_step = [&](int buf, const SA& xa, const SB& xb) {
        char* base = smem + buf * S::BUF;
        xa.store(AKC, base, S::A_PLANE, S::ROWB, tid);
        xb.store(BKC, base + AR::NPLANE * S::A_PLANE, S::B_PLANE, S::ROWB, tid);
    };
    auto compute = [&](int buf) {
        const char* la = smem + buf * S::BUF;
        tile_compute<AR, S, BK>(acc, la, la + AR::NPLANE * S::A_PLANE, wm, wn, lane);
    };

    if constexpr (STAGES == 1) {
        load_step(0, sa[0], sb[0]);
        __syncthreads();  // previous user of smem is done
        store_step(0, sa[0], sb[0]);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const int buf = kt & 1;
            if (kt + 1 < nk) load_step(kt + 1, sa[0], sb[0]);
            compute(buf);
            if (kt + 1 < nk) store_step(buf ^ 1, sa[0], sb[0]);
            __syncthreads();
        }
    } else {
        // staging set of k-block u is u % 2; LDS buffer of k-block u is u % 2
        load_step(0, sa[0], sb[0]);
        if (nk > 1) load_step(1, sa[1], sb[1]);
        __syncthreads();
        store_step(0, sa[0], sb[0]);
        __syncthreads();
        for (int kt = 0; kt < nk; kt += 2) {
            // even k-block kt: set 0 is in LDS buffer 0, set 1 holds kt+1
            if (kt + 2 < nk) load_step(kt + 2, sa[0], sb[0]);
            compute(0);
            if (kt + 1 < nk) store_step(1, sa[1], sb[1]);
            __syncthreads();
            if (kt + 1 >= nk) break;
            // odd k-block kt+1: set 1 is in LDS buffer 1, set 0 holds kt+2
            if (kt + 3 < nk) load_step(kt + 3, sa[1], sb[1]);
            compute(1);
            if (kt + 2 < nk) store_step(0, sa[0], sb[0]);
            __syncthreads();
        }
    }

    // fused alpha/beta epilogue
    const int rb = j.i0 + wm * S::WTM + 4 * (lane >> 4), cb = j.j0 + wn * S::WTN + (lane & 15);
    if constexpr (PRE) {
        if (j.beta0 == 0) {
#pragma unroll
            for (int fm = 0; fm < S::FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < S::FN; ++fn)
                    store_frag_old<AR>(C, j.ldc, rb + fm * 16, cb + fn * 16, acc[fm][fn], j.s, c_fast, j.m, j.n,
                                       old[fm][fn]);
            return;
        }
    }
    store_tile<AR>(C, j.ldc, acc, j.s, j.beta0 != 0, c_fast, j.m, j.n, [&](int fm, int fn, int& r0, int& c) {
        r0 = rb + fm * 16;
        c = cb + fn * 16;
    });
}

__device__ __forceinline__ TileJob make_job(const KParams& p, int inst, int i0, int j0)
{
    TileJob j;
    j.A = static_cast<const char*>(p.A[inst]);
    j.B = static_cast<const char*>(p.B[inst]);
    j.C = p.C[inst];
    j.lda = p.lda;
    j.ldb = p.ldb;
    j.ldc = p.ldc;
    j.m = p.m;
    j.n = p.n;
    j.k = p.k;
    j.i0 = i0;
    j.j0 = j0;
    j.s = p.s;
    j.beta0 = p.beta0;
    j.akc = 0;
    j.bkc = 0;
    return j;
}

// Batched persistent kernel: workgroups walk the (instance, tile) work ids.
template <class AR, bool AKC, bool BKC, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int STAGES = 1>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
tile_gemm(const KParams p)
{
    using S = TileShape<AR, AKC, BKC, BM, BN, BK, WAVES_M, WAVES_N>;
    __shared__ __attribute__((aligned(16))) char smem[S::LDS_BYTES];
    const int tiles_per_inst = p.tiles_m * p.tiles_n;
    for (long t = blockIdx.x; t < p.total; t += gridDim.x) {
        const long w = xcd_remap(t, p.total);
        const int inst = (int)(w / tiles_per_inst);
        const int rem = (int)(w % tiles_per_inst);
        const TileJob j = make_job(p, inst, (rem % p.tiles_m) * BM, (rem / p.tiles_m) * BN);
        tile_job<AR, AKC ? 1 : 0, BKC ? 1 : 0, BM, BN, BK, WAVES_M, WAVES_N, STAGES>(j, smem);
    }
}

// ---------------------------------------------------------------------------
// Direct (LDS-free) tiles: one wave per WTM x WTN block.
// ---------------------------------------------------------------------------

// Load the lane's operand fragment for one MFMA k-step straight from global:
// row `r` of the operand view, input elements k = kb + g*KL + s, s < KL,
// KL = KSTEP/4.  Returns NPLANE 16-byte plane chunks.
template <class AR, bool KC>
__device__ __forceinline__ void direct_frag(const char* base, long sr, long sk, int r, int rows,
                                            int kb, int kvalid, bool vec, u32x4* out)
{
    constexpr int E = AR::E;
    constexpr int KL = AR::KSTEP / 4;          // input elements per lane per k-step
    constexpr int NIN = KL * E / 16;           // 16-byte input words (1, or 2 for bf16)
    using RT = typename Raw<E>::T;
    u32x4 in[NIN];
    const int g = threadIdx.x & 63;
    const int k0 = kb + (g >> 4) * KL;
    if (KC && vec && r < rows && k0 + KL <= kvalid) {
#pragma unroll
        for (int w = 0; w < NIN; ++w)
            in[w] = *gbl(reinterpret_cast<const u32x4*>(base + ((long)r * sr + k0) * E + 16 * w));
    } else {
        RT vals[KL];
#pragma unroll
        for (int s = 0; s < KL; ++s) {
            bool ok = (r < rows) && (k0 + s < kvalid);
            vals[s] = ok ? load_elem<E>(base, (long)r * sr + (long)(k0 + s) * sk) : RT(0);
        }
#pragma unroll
        for (int w = 0; w < NIN; ++w) __builtin_memcpy(&in[w], &vals[w * 16 / E], 16);
    }
    if constexpr (AR::LDS_E == E) {
        out[0] = in[0];
    } else {
        // bf16 split: two 16-byte input words -> one 16-byte chunk per plane
        u32x4 lo[AR::NPLANE], hi[AR::NPLANE];
        AR::to_planes(in[0], lo);
        AR::to_planes(in[NIN - 1], hi);
#pragma unroll
        for (int pl = 0; pl < AR::NPLANE; ++pl) out[pl] = u32x4{lo[pl][0], lo[pl][1], hi[pl][0], hi[pl][1]};
    }
}

// One wave computes the WTM x WTN block of C at (j.i0, j.j0).
template <class AR, bool AKC, bool BKC, int WTM, int WTN>
__device__ __forceinline__ void direct_job(const TileJob& j)
{
    using Out = typename AR::Out;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    const int lane = threadIdx.x & 63;
    const long sra = AKC ? j.lda : 1, ska = AKC ? 1 : j.lda;
    const long srb = BKC ? j.ldb : 1, skb = BKC ? 1 : j.ldb;
    const bool a_vec = ((j.lda * AR::E) % 16) == 0 && aligned16(j.A);
    const bool b_vec = ((j.ldb * AR::E) % 16) == 0 && aligned16(j.B);

    typename AR::Acc acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = typename AR::Acc{};

    // old C values (beta != 0) read before the k-loop
    Out* C = static_cast<Out*>(j.C);
    const bool c_fast = j.i0 + WTM <= j.m && j.j0 + WTN <= j.n && ((j.ldc * (long)sizeof(Out)) % 16) == 0 &&
                        aligned16(C);
    Out old[FM][FN][4];
    if (j.beta0 == 0) {
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
                load_old_frag<AR>(C, j.ldc, j.i0 + fm * 16 + 4 * (lane >> 4), j.j0 + fn * 16 + (lane & 15), c_fast,
                                  j.m, j.n, old[fm][fn]);
    }

    for (int kb = 0; kb < j.k; kb += AR::KSTEP) {
        u32x4 af[FM][AR::NPLANE], bfg[FN][AR::NPLANE];
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
            direct_frag<AR, AKC>(j.A, sra, ska, j.i0 + fm * 16 + frag_row<AR>(lane & 15), j.m, kb, j.k, a_vec,
                                 af[fm]);
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
            direct_frag<AR, BKC>(j.B, srb, skb, j.j0 + fn * 16 + (lane & 15), j.n, kb, j.k, b_vec, bfg[fn]);
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) AR::mma(acc[fm][fn], af[fm], bfg[fn]);
    }

    const int rb = j.i0 + 4 * (lane >> 4), cb = j.j0 + (lane & 15);
    if (j.beta0 == 0) {
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
                store_frag_old<AR>(C, j.ldc, rb + fm * 16, cb + fn * 16, acc[fm][fn], j.s, c_fast, j.m, j.n,
                                   old[fm][fn]);
    } else {
        store_tile<AR>(C, j.ldc, acc, j.s, true, c_fast, j.m, j.n, [&](int fm, int fn, int& r0, int& c) {
            r0 = rb + fm * 16;
            c = cb + fn * 16;
        });
    }
}

template <class AR, bool AKC, bool BKC, int WTM, int WTN>
__global__ void __launch_bounds__(256)
direct_gemm(const KParams p)
{
    const int per_inst = p.tiles_m * p.tiles_n;
    const long nwaves = (long)gridDim.x * 4;
    // the wave index as a scalar, so the work id, and the instance pointers
    // read with it, are known to be the same in every lane
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (long t = (long)blockIdx.x * 4 + wave; t < p.total; t += nwaves) {
        const int inst = (int)(t / per_inst);
        const int rem = (int)(t % per_inst);
        const TileJob j = make_job(p, inst, (rem % p.tiles_m) * WTM, (rem / p.tiles_m) * WTN);
        direct_job<AR, AKC, BKC, WTM, WTN>(j);
    }
}

}  // namespace ssb
