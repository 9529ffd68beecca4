// ozaki.hip -- fp64 batched GEMM emulated on the int8 MFMA units (Ozaki
// scheme with integer slices: Ozaki, Ogita, Oishi, Rump, Numer. Algorithms
// 59 (2012); int8 form after Ootomo, Ozaki, Yokota, IJHPCA 2024).
//
// For every instance:
//   1. e_i = exponent with max_k |op(A)(i,k)| < 2^e_i  (row i of op(A)),
//      f_j likewise for column j of op(B)               [ozaki_exp_* kernels]
//   2. A'(i,k) = op(A)(i,k) * 2^-e_i in (-1, 1) is cut into s signed 7-bit
//      slices:  r = A' ; repeat s times { r *= 128; a_p = trunc(r); r -= a_p }
//      so A' = sum_p a_p 2^(-7p) + O(2^(-7s)), |a_p| <= 127; same for B.
//      Every step is exact in fp64.  Slices are stored with the operand's
//      own orientation and a compact leading dimension [ozaki_slice kernel]
//   3. for t = s+1 down to 2:  D_t = sum_{p+q=t} a_p b_q  -- exact int32
//      products on v_mfma_i32_16x16x64_i8 -- and D += D_t * 2^(-7t) in fp64
//      (smallest terms first)                                [ozaki_gemm]
//   4. C = fma(D * 2^(e_i+f_j), alpha, fl(beta * C))  (the reference's
//      epilogue rounding order)
// The triangle p + q <= s + 1 keeps s(s+1)/2 of the s^2 slice products; the
// dropped ones are O(2^(-7(s+1))) relative to the row/column maxima.
// int32 exactness: one D_t sums at most s*kc*127^2 < 2^31 for a K chunk kc
// of at most 8192; longer K is folded chunk by chunk.
// Rows or columns holding inf/nan get NaN in every output they touch.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>

#include "gemm_kernels.hpp"
#include "api_internal.hpp"
#include "launch.hpp"
#include "occupancy.hpp"

namespace ssb {

struct OzParams {
    const void* const* A;  // caller's fp64 operands
    const void* const* B;
    void* const* C;
    long lda, ldb, ldc;
    int m, n, k, batch, s;
    int ta, tb;
    // workspace
    int8_t* wa;      // A slices: [inst][p][cols_sa][ld_sa]
    int8_t* wb;      // B slices
    int32_t* ea;     // [inst][m] exponents (INT_MIN marks a non-finite row)
    int32_t* eb;     // [inst][n]
    long ld_sa, cols_sa, ld_sb, cols_sb;  // slice matrices: storage ld and column count
    Scalars sc;
    int beta0;
    int tiles_m, tiles_n;
    long total;
};

constexpr int OZ_NONFINITE = INT32_MIN;
constexpr int OZ_KCHUNK = 8192;

// ---------------------------------------------------------------------------
// exponents
// ---------------------------------------------------------------------------
__device__ __forceinline__ int exp_of_max(double mx, bool nonfinite)
{
    if (nonfinite) return OZ_NONFINITE;
    if (mx == 0.0) return 0;
    int e;
    (void)frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1)  ->  mx < 2^e
    return e;
}

// Reduce over the contiguous dimension: one wave per storage column.
// out[inst * ncols + c] = exponent of max_r |X(r, c)|, r < nrows.
__global__ void __launch_bounds__(256) ozaki_exp_cols(const void* const* X, long ld, int nrows, int ncols,
                                                      int32_t* out)
{
    const int inst = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + wave;
    if (c >= ncols) return;
    const double* x = static_cast<const double*>(X[inst]) + (long)c * ld;
    double mx = 0.0;
    bool bad = false;
    for (int r = lane; r < nrows; r += 64) {
        double v = fabs(*gbl(x + r));
        bad |= !isfinite(v);
        mx = fmax(mx, v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off));
        bad |= __shfl_xor((int)bad, off) != 0;
    }
    if (lane == 0) out[(long)inst * ncols + c] = exp_of_max(mx, bad);
}

// Reduce over the strided dimension: one thread per storage row, columns
// split over the 4 waves of the block.
__global__ void __launch_bounds__(256) ozaki_exp_rows(const void* const* X, long ld, int nrows, int ncols,
                                                      int32_t* out)
{
    __shared__ double smx[4][64];
    __shared__ int sbad[4][64];
    const int inst = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    const double* x = static_cast<const double*>(X[inst]);
    double mx = 0.0;
    bool bad = false;
    if (r < nrows) {
        for (int c = wave; c < ncols; c += 4) {
            double v = fabs(*gbl(x + r + (long)c * ld));
            bad |= !isfinite(v);
            mx = fmax(mx, v);
        }
    }
    smx[wave][lane] = mx;
    sbad[wave][lane] = bad;
    __syncthreads();
    if (wave == 0 && r < nrows) {
        mx = fmax(fmax(smx[0][lane], smx[1][lane]), fmax(smx[2][lane], smx[3][lane]));
        bad = sbad[0][lane] | sbad[1][lane] | sbad[2][lane] | sbad[3][lane];
        out[(long)inst * nrows + r] = exp_of_max(mx, bad);
    }
}

// ---------------------------------------------------------------------------
// slices: elementwise over the storage of one operand
// ---------------------------------------------------------------------------
// exp_by_col: the exponent of element (r, c) is e[c] (else e[r]).
__global__ void __launch_bounds__(256) ozaki_slice(const void* const* X, long ld, int nrows, int ncols,
                                                   const int32_t* e, int exp_by_col, int s, int8_t* w,
                                                   long ldw)
{
    const int inst = blockIdx.z;
    const int c = blockIdx.y;
    const int r0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (r0 >= ldw) return;
    const double* x = static_cast<const double*>(X[inst]) + (long)c * ld;
    const int32_t* ei = e + (long)inst * (exp_by_col ? ncols : nrows);
    const long plane = ldw * ncols;
    int8_t* out = w + (long)inst * s * plane + (long)c * ldw + r0;
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = r0 + q;
        v[q] = 0.0;
        if (r < nrows) {
            const int ex = exp_by_col ? ei[c] : ei[r];
            v[q] = (ex == OZ_NONFINITE) ? 0.0 : ldexp(*gbl(x + r), -ex);
        }
    }
    for (int p = 0; p < s; ++p) {
        uint32_t word = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] *= 128.0;
            const double t = trunc(v[q]);
            v[q] -= t;
            word |= ((uint32_t)(int32_t)t & 0xffu) << (8 * q);
        }
        *gbl(reinterpret_cast<uint32_t*>(out + p * plane)) = word;
    }
}

// ---------------------------------------------------------------------------
// multi-product int8 GEMM with fp64 fold
// ---------------------------------------------------------------------------
template <bool AKC, bool BKC, int BM, int BN, int BK>
__global__ void __launch_bounds__(256) ozaki_gemm(const OzParams p)
{
    using AR = ArithI8;
    constexpr int NT = 256, WAVES_M = 2;
    constexpr int WTM = BM / 2, WTN = BN / 2, FM = WTM / 16, FN = WTN / 16;
    constexpr int ROWB = BK + 16;
    constexpr int A_PLANE = BM * ROWB, B_PLANE = BN * ROWB, BUF = A_PLANE + B_PLANE;
    __shared__ __attribute__((aligned(16))) char smem[2 * BUF];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;
    // slice-operand views (int8, compact leading dimensions)
    const long sra = AKC ? p.ld_sa : 1, ska = AKC ? 1 : p.ld_sa;
    const long srb = BKC ? p.ld_sb : 1, skb = BKC ? 1 : p.ld_sb;
    const long plane_a = p.ld_sa * p.cols_sa, plane_b = p.ld_sb * p.cols_sb;
    const int tiles_per_inst = p.tiles_m * p.tiles_n;
    const int s = p.s;

    using SA = Stager<AR, AKC ? 1 : 0, BM, BK, NT>;
    using SB = Stager<AR, BKC ? 1 : 0, BN, BK, NT>;
    SA sa;
    SB sb;

    for (long t = blockIdx.x; t < p.total; t += gridDim.x) {
        const long w = xcd_remap(t, p.total);
        const int inst = (int)(w / tiles_per_inst);
        const int rem = (int)(w % tiles_per_inst);
        const int i0 = (rem % p.tiles_m) * BM;
        const int j0 = (rem / p.tiles_m) * BN;
        const int rows_a = min(BM, p.m - i0), rows_b = min(BN, p.n - j0);
        const char* Ai = reinterpret_cast<const char*>(p.wa) + (long)inst * s * plane_a + (long)i0 * sra;
        const char* Bi = reinterpret_cast<const char*>(p.wb) + (long)inst * s * plane_b + (long)j0 * srb;
        const bool a_vec = (p.ld_sa % SA::word_bytes(AKC)) == 0, b_vec = (p.ld_sb % SB::word_bytes(BKC)) == 0;
        const bool a_fast = a_vec && rows_a == BM, b_fast = b_vec && rows_b == BN;

        f64x4 D[FM][FN];
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) D[fm][fn] = f64x4{0.0, 0.0, 0.0, 0.0};

        for (int kc0 = 0; kc0 < p.k; kc0 += OZ_KCHUNK) {
            const int kc = min(OZ_KCHUNK, p.k - kc0);
            const int nk = (kc + BK - 1) / BK;
            for (int tt = s + 1; tt >= 2; --tt) {
                const int plo = max(1, tt - s), phi = min(s, tt - 1);
                const int nsteps = (phi - plo + 1) * nk;
                i32x4 acc[FM][FN];
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = i32x4{0, 0, 0, 0};

                auto load_step = [&](int u) {
                    const int pa = plo + u / nk, pb = tt - pa;
                    const int k0 = kc0 + (u % nk) * BK;
                    const int kv = p.k - k0;
                    const char* ak = Ai + (long)(pa - 1) * plane_a + (long)k0 * ska;
                    const char* bk = Bi + (long)(pb - 1) * plane_b + (long)k0 * skb;
                    const int kv_c = min(kv, kc0 + kc - k0);
                    if (a_fast && kv_c >= BK) sa.load_fast(AKC, ak, sra, ska, tid);
                    else sa.load_guarded(AKC, ak, sra, ska, rows_a, kv_c, a_vec, tid);
                    if (b_fast && kv_c >= BK) sb.load_fast(BKC, bk, srb, skb, tid);
                    else sb.load_guarded(BKC, bk, srb, skb, rows_b, kv_c, b_vec, tid);
                };
                auto store_step = [&](int buf) {
                    char* base = smem + buf * BUF;
                    sa.store(AKC, base, A_PLANE, ROWB, tid);
                    sb.store(BKC, base + A_PLANE, B_PLANE, ROWB, tid);
                };

                load_step(0);
                __syncthreads();
                store_step(0);
                __syncthreads();
                for (int u = 0; u < nsteps; ++u) {
                    const int buf = u & 1;
                    if (u + 1 < nsteps) load_step(u + 1);
                    const char* la = smem + buf * BUF;
                    const char* lb = la + A_PLANE;
#pragma unroll
                    for (int ks = 0; ks < BK / AR::KSTEP; ++ks) {
                        const int kbyte = ks * AR::KSTEP + (lane >> 4) * 16;
                        u32x4 af[FM], bfg[FN];
#pragma unroll
                        for (int fm = 0; fm < FM; ++fm)
                            af[fm] = *reinterpret_cast<const u32x4*>(la + (wm * WTM + fm * 16 + (lane & 15)) * ROWB + kbyte);
#pragma unroll
                        for (int fn = 0; fn < FN; ++fn)
                            bfg[fn] = *reinterpret_cast<const u32x4*>(lb + (wn * WTN + fn * 16 + (lane & 15)) * ROWB + kbyte);
#pragma unroll
                        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                            for (int fn = 0; fn < FN; ++fn) AR::mma(acc[fm][fn], &af[fm], &bfg[fn]);
                    }
                    if (u + 1 < nsteps) store_step(buf ^ 1);
                    __syncthreads();
                }
                // fold: D += D_t * 2^(-7 t)
                const double scale = ldexp(1.0, -7 * tt);
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn) {
                        int32_t a4[4];
                        __builtin_memcpy(a4, &acc[fm][fn], 16);
                        double d4[4];
                        __builtin_memcpy(d4, &D[fm][fn], 32);
#pragma unroll
                        for (int v = 0; v < 4; ++v) d4[v] = __builtin_fma((double)a4[v], scale, d4[v]);
                        __builtin_memcpy(&D[fm][fn], d4, 32);
                    }
            }
        }

        // epilogue: C = fma(D * 2^(e_i + f_j), alpha, fl(beta * C))
        double* C = static_cast<double*>(p.C[inst]);
        const int32_t* ea = p.ea + (long)inst * p.m;
        const int32_t* eb = p.eb + (long)inst * p.n;
        const bool beta0 = p.beta0 != 0;
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) {
                const int r0 = i0 + wm * WTM + fm * 16 + 4 * (lane >> 4);
                const int c = j0 + wn * WTN + fn * 16 + (lane & 15);
                if (c >= p.n) continue;
                const int fj = eb[c];
                double d4[4];
                __builtin_memcpy(d4, &D[fm][fn], 32);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = r0 + v;
                    if (r >= p.m) break;
                    const int ei = ea[r];
                    double val = (ei == OZ_NONFINITE || fj == OZ_NONFINITE) ? __builtin_nan("") : ldexp(d4[v], ei + fj);
                    double* dst = C + (long)c * p.ldc + r;
                    double t0 = beta0 ? 0.0 : (*gbl(dst)) * p.sc.bd;
                    *gbl(dst) = __builtin_fma(val, p.sc.ad, t0);
                }
            }
    }
}

// ---------------------------------------------------------------------------
// fused form for small instances (m, n, k <= 32): one wave per instance,
// nothing in HBM but A, B and C, and no LDS.
//
// Slicing: lane l owns row l % 32 of op(A) and column l % 32 of op(B) over
// the k half kh = l / 32 (16 doubles each); the row / column maximum joins
// the two halves with one lane exchange.  Each slice of the 16 values is
// one 16-byte register chunk.
//
// Products: v_mfma_i32_32x32x32_i8 takes its A operand as lane l -> row
// l % 32, 16 consecutive k of half l / 32, and B likewise with columns --
// exactly the chunk the lane sliced -- so one instruction is the whole
// 32x32x32 product of one slice pair, straight from registers.  Its result
// is lane l -> column l % 32, accumulator v -> row 8*(v/4) + 4*(l/32) + v%4.
// D_t = sum of the pairs of order t is exact in int32 (9 * 32 * 127^2 <
// 2^31); the fold (smallest order first), the exponents and the epilogue
// are ozaki_gemm's, so the result is bit-identical to the general path.
// ---------------------------------------------------------------------------
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));

// The 7-bit digits of a 28-bit value, one per byte: byte 3 the most
// significant digit, byte 0 the least.
__device__ __forceinline__ uint32_t oz_spread7(uint32_t h)
{
    return (h & 0x7fu) | ((h << 1) & 0x7f00u) | ((h << 2) & 0x7f0000u) | ((h << 3) & 0x7f000000u);
}

// Slices of 16 values (v = x * 2^-e, |v| < 1) into S 16-byte chunks: slice
// p holds, for every value, its p-th base-128 digit with the value's sign
// (the digits the sequence "v *= 128; d = trunc(v); v -= d" gives, as in
// ozaki_slice).  Done in integers: H = floor(|v| 2^28) carries digits 1-4,
// L = floor(frac(|v| 2^28) 2^28) digits 5-8 and floor(frac(|v| 2^56) 2^7)
// digit 9 (every step exact in fp64); the 7-bit fields are spread to bytes
// and transposed across 4 values; a negative value's bytes are negated at
// once (bytes below 128: -d = ((0x80 - d) ^ 0x80) without borrows).
template <int S>
__device__ __forceinline__ void oz_slice_regs(const double* x, int e, u32x4 (&sl)[S])
{
    static_assert(S >= 1 && S <= 9, "1 to 9 slices");
    uint32_t H[16], L[16], N[16], neg[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const double a = (e == OZ_NONFINITE) ? 0.0 : ldexp(fabs(x[i]), -e);
        neg[i >> 2] |= (x[i] < 0.0 ? 0xffu : 0u) << (8 * (i & 3));
        const double y1 = ldexp(a, 28);
        H[i] = (uint32_t)y1;
        if constexpr (S > 4) {
            const double y2 = ldexp(y1 - (double)H[i], 28);
            L[i] = (uint32_t)y2;
            if constexpr (S > 8) N[i] = (uint32_t)ldexp(y2 - (double)L[i], 7);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t w[9];
        {
            const uint32_t in[4] = {oz_spread7(H[4 * j]), oz_spread7(H[4 * j + 1]), oz_spread7(H[4 * j + 2]),
                                    oz_spread7(H[4 * j + 3])};
            uint32_t out[4];
            transpose4x4_bytes(in, out);  // out[c]: byte c of the four values
            w[0] = out[3];
            w[1] = out[2];
            w[2] = out[1];
            w[3] = out[0];
        }
        if constexpr (S > 4) {
            const uint32_t in[4] = {oz_spread7(L[4 * j]), oz_spread7(L[4 * j + 1]), oz_spread7(L[4 * j + 2]),
                                    oz_spread7(L[4 * j + 3])};
            uint32_t out[4];
            transpose4x4_bytes(in, out);
            w[4] = out[3];
            w[5] = out[2];
            w[6] = out[1];
            w[7] = out[0];
        }
        if constexpr (S > 8) w[8] = N[4 * j] | (N[4 * j + 1] << 8) | (N[4 * j + 2] << 16) | (N[4 * j + 3] << 24);
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const uint32_t nw = (0x80808080u - w[p]) ^ 0x80808080u;
            sl[p][j] = (neg[j] & nw) | (~neg[j] & w[p]);
        }
    }
}

// Exponent of the max |.| over both k halves (lanes l and l ^ 32).
__device__ __forceinline__ int oz_exp16(const double* x)
{
    double mx = 0.0;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const double a = fabs(x[i]);
        bad |= !isfinite(a);
        mx = fmax(mx, a);
    }
    mx = fmax(mx, __shfl_xor(mx, 32));
    bad |= __shfl_xor((int)bad, 32) != 0;
    return exp_of_max(mx, bad);
}

// Where a lane's 16 doubles sit in one operand: row `own` of op(A) (column
// `own` of op(B)) over the k half kh.  Indices are clamped into the matrix,
// so every load is in bounds and unconditional; elements outside the
// instance are zeroed afterwards (only when it is not a full 32 x 32 block).
struct OzLane {
    uint32_t base, step;  // element offset of k = 16 kh and per k (both < 2^32: a 32x32 block of one instance)
    uint32_t kvalid;      // elements of the half inside the instance (0..16)
    bool row_ok;
};
__device__ __forceinline__ OzLane oz_lane(long ld, bool kcontig, int rows, int k, int own, int kh)
{
    OzLane L;
    const uint32_t r = (uint32_t)min(own, rows - 1);
    const uint32_t k0 = (uint32_t)min(16 * kh, k - 1);
    L.step = kcontig ? 1u : (uint32_t)ld;
    L.base = kcontig ? r * (uint32_t)ld + k0 : r + k0 * (uint32_t)ld;
    L.kvalid = (uint32_t)max(0, min(16, k - 16 * kh));
    L.row_ok = own < rows;
    return L;
}
// A pointer read from a pointer array at a wave-uniform index, as a scalar
// (the compiler cannot use a scalar load for it: the kernel's stores might
// alias the array), so element addresses are scalar base + 32-bit offset.
template <class T>
__device__ __forceinline__ T* wave_uniform(T* ptr)
{
    const uint64_t v = reinterpret_cast<uint64_t>(ptr);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// FULL: the instance is a whole 32 x 32 x 32 block (no element to zero).
template <bool FULL>
__device__ __forceinline__ void oz_small_load(const void* X, const OzLane& L, double (&x)[16])
{
    const double* P = wave_uniform(static_cast<const double*>(X));
    if constexpr (FULL) {
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = *gbl(P + (L.base + (uint32_t)q * L.step));
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t qq = (uint32_t)q < L.kvalid ? (uint32_t)q : 0u;  // stays inside the matrix
            const double v = *gbl(P + (L.base + qq * L.step));
            x[q] = (L.row_ok && (uint32_t)q < L.kvalid) ? v : 0.0;
        }
    }
}

// Persistent: wave w of the grid computes instances w, w + W, w + 2W, ...
// (W = waves in the grid), loading each instance's operands straight into
// registers.  (Copying the next instance global -> LDS while computing the
// current one was measured slower: the extra LDS reads and waits cost more
// than the load latency they hide at two waves per SIMD.)
template <int S, bool FULL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) ozaki_small(const OzParams p)
{
    const int lane = threadIdx.x & 63;
    // wave index as a scalar: instance numbers, and so the operand pointers
    // read from the pointer arrays, are then known to be wave-uniform
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long W = (long)gridDim.x * 4;
    long inst = (long)blockIdx.x * 4 + wave;
    if (inst >= p.batch) return;
    const int own = lane & 31, kh = lane >> 5;
    const bool beta0 = p.beta0 != 0;
    // op(A) row own is K-contiguous when A is transposed; op(B) column own
    // when B is not
    const OzLane LA = oz_lane(p.lda, p.ta != 0, p.m, p.k, own, kh);
    const OzLane LB = oz_lane(p.ldb, p.tb == 0, p.n, p.k, own, kh);

    for (; inst < p.batch; inst += W) {
        double xa[16], xb[16];
        oz_small_load<FULL>(p.A[inst], LA, xa);
        oz_small_load<FULL>(p.B[inst], LB, xb);
        u32x4 sa[S], sb[S];
        const int ea = oz_exp16(xa);
        oz_slice_regs<S>(xa, ea, sa);
        const int eb = oz_exp16(xb);
        oz_slice_regs<S>(xb, eb, sb);

        double D[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) D[v] = 0.0;
#pragma unroll
        for (int t = S + 1; t >= 2; --t) {
            const int plo = t - S > 1 ? t - S : 1, phi = S < t - 1 ? S : t - 1;
            i32x16 acc = {};
#pragma unroll
            for (int pa = plo; pa <= phi; ++pa)
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8((i32x4)sa[pa - 1], (i32x4)sb[t - pa - 1], acc, 0, 0, 0);
            const double scale = ldexp(1.0, -7 * t);
#pragma unroll
            for (int v = 0; v < 16; ++v) D[v] = __builtin_fma((double)acc[v], scale, D[v]);
            // fold each order here: left alone, the compiler sinks the folds
            // into the epilogue and keeps S accumulators of 16 registers live
#pragma unroll
            for (int v = 0; v < 16; ++v) __asm__ volatile("" : "+v"(D[v]));
        }

        // epilogue: C = fma(D * 2^(e_i + f_j), alpha, fl(beta * C)); the
        // lane's 4 consecutive rows of its column go out as two 16-byte
        // stores when they are whole and aligned
        double* C = wave_uniform(static_cast<double*>(p.C[inst]));
        const bool vec = (p.ldc % 2) == 0 && aligned16(C);
        const int c = own;
        // row exponents, shuffled while every lane is active (lane `row`
        // holds that row's)
        int eir[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) eir[v] = __shfl(ea, 8 * (v >> 2) + 4 * kh + (v & 3));
        if (c < p.n) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row0 = 8 * q + 4 * kh;
                double val[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int ei = eir[4 * q + v];
                    val[v] = (ei == OZ_NONFINITE || eb == OZ_NONFINITE) ? __builtin_nan("")
                                                                        : ldexp(D[4 * q + v], ei + eb);
                }
                double* dst = C + (long)c * p.ldc + row0;
                if (vec && row0 + 4 <= p.m) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        f64x2 old = {0.0, 0.0};
                        if (!beta0) old = *gbl(reinterpret_cast<const f64x2*>(dst + 2 * h));
                        f64x2 res;
                        res[0] = __builtin_fma(val[2 * h], p.sc.ad, beta0 ? 0.0 : old[0] * p.sc.bd);
                        res[1] = __builtin_fma(val[2 * h + 1], p.sc.ad, beta0 ? 0.0 : old[1] * p.sc.bd);
                        *gbl(reinterpret_cast<f64x2*>(dst + 2 * h)) = res;
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        if (row0 + v < p.m) {
                            const double t0 = beta0 ? 0.0 : (*gbl(dst + v)) * p.sc.bd;
                            *gbl(dst + v) = __builtin_fma(val[v], p.sc.ad, t0);
                        }
                    }
                }
            }
        }
    }
}

template <int S, bool FULL>
static hipError_t run_ozaki_small_k(const OzParams& p, hipStream_t st)
{
    const int occ = kernel_occupancy(ozaki_small<S, FULL>, 256);
    const long blocks = ((long)p.batch + 3) / 4;
    long grid = (long)occ * device_cu_count();
    if (const char* e = getenv("SSBLAS_MAX_GRID")) {
        const long cap = atol(e);
        if (cap >= 1 && cap < grid) grid = cap;
    }
    if (blocks < grid) grid = blocks;
    hipLaunchKernelGGL((ozaki_small<S, FULL>), dim3((unsigned)grid), dim3(256), 0, st, p);
    return hipGetLastError();
}

template <int S>
static hipError_t run_ozaki_small(const OzParams& p, hipStream_t st)
{
    if (p.m == 32 && p.n == 32 && p.k == 32) return run_ozaki_small_k<S, true>(p, st);
    return run_ozaki_small_k<S, false>(p, st);
}

static hipError_t ozaki_small_dispatch(const OzParams& p, hipStream_t st)
{
    switch (p.s) {
    case 3: return run_ozaki_small<3>(p, st);
    case 4: return run_ozaki_small<4>(p, st);
    case 5: return run_ozaki_small<5>(p, st);
    case 6: return run_ozaki_small<6>(p, st);
    case 7: return run_ozaki_small<7>(p, st);
    case 8: return run_ozaki_small<8>(p, st);
    default: return run_ozaki_small<9>(p, st);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <bool AKC, bool BKC>
static hipError_t run_ozaki_gemm(OzParams p, hipStream_t st)
{
    constexpr int BM = 64, BN = 64, BK = 64;
    auto kern = ozaki_gemm<AKC, BKC, BM, BN, BK>;
    p.tiles_m = (p.m + BM - 1) / BM;
    p.tiles_n = (p.n + BN - 1) / BN;
    p.total = (long)p.batch * p.tiles_m * p.tiles_n;
    const int occ = kernel_occupancy(kern, 256);
    long slots = (long)occ * device_cu_count();
    long grid = p.total <= slots ? p.total : (slots / 8) * 8;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, p);
    return hipGetLastError();
}

// Exponents + slices of one operand.  Storage is nrows x ncols (column
// major); exponents are per storage column when exp_by_col.
static hipError_t prep_operand(const void* const* X, long ld, int nrows, int ncols, int exp_by_col, int s,
                               int batch, int32_t* e, int8_t* w, long ldw, hipStream_t st)
{
    if (exp_by_col) {
        hipLaunchKernelGGL(ozaki_exp_cols, dim3((ncols + 3) / 4, batch), dim3(256), 0, st, X, ld, nrows, ncols, e);
    } else {
        hipLaunchKernelGGL(ozaki_exp_rows, dim3((nrows + 63) / 64, batch), dim3(256), 0, st, X, ld, nrows, ncols, e);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(ozaki_slice, dim3((unsigned)((ldw / 4 + 255) / 256), ncols, batch), dim3(256), 0, st, X, ld,
                       nrows, ncols, e, exp_by_col, s, w, ldw);
    return hipGetLastError();
}

hipError_t ozaki_run(const GemmProblem& g, int s, WsGet ws_get, hipStream_t st)
{
    const int m = g.m, n = g.n, k = g.k;
    if (ozaki_fused_shape(g) && s >= 3 && s <= 9) {
        // fused single-wave form: no workspace
        OzParams p{};
        p.A = g.A;
        p.B = g.B;
        p.C = g.C;
        p.lda = g.lda;
        p.ldb = g.ldb;
        p.ldc = g.ldc;
        p.m = m;
        p.n = n;
        p.k = k;
        p.batch = g.batch;
        p.s = s;
        p.ta = g.ta;
        p.tb = g.tb;
        p.sc.ad = g.ad;
        p.sc.bd = g.bd;
        p.beta0 = g.beta0 ? 1 : 0;
        return ozaki_small_dispatch(p, st);
    }
    const OzakiLayout L = ozaki_layout(g, s);
    hipError_t err = hipSuccess;
    char* ws = static_cast<char*>(ws_get(L.total, &err));
    if (ws == nullptr) return err == hipSuccess ? hipErrorOutOfMemory : err;
    const long ra = g.ta ? k : m, ca = g.ta ? m : k;  // storage of A
    const long rb = g.tb ? n : k, cb = g.tb ? k : n;  // storage of B
    const long ld_sa = L.ld_sa, ld_sb = L.ld_sb;

    OzParams p;
    p.A = g.A;
    p.B = g.B;
    p.C = g.C;
    p.lda = g.lda;
    p.ldb = g.ldb;
    p.ldc = g.ldc;
    p.m = m;
    p.n = n;
    p.k = k;
    p.batch = g.batch;
    p.s = s;
    p.ta = g.ta;
    p.tb = g.tb;
    p.wa = reinterpret_cast<int8_t*>(ws);
    p.wb = reinterpret_cast<int8_t*>(ws + L.off_b);
    p.ea = reinterpret_cast<int32_t*>(ws + L.off_e);
    p.eb = p.ea + (size_t)g.batch * m;
    p.ld_sa = ld_sa;
    p.cols_sa = ca;
    p.ld_sb = ld_sb;
    p.cols_sb = cb;
    p.sc.af = g.af;
    p.sc.bf = g.bf;
    p.sc.ad = g.ad;
    p.sc.bd = g.bd;
    p.sc.ai = g.ai;
    p.sc.bi = g.bi;
    p.beta0 = g.beta0 ? 1 : 0;

    // op(A) row i: storage row i (N) -> reduce along the strided dim; storage column i (T)
    err = prep_operand(g.A, g.lda, (int)ra, (int)ca, g.ta ? 1 : 0, s, g.batch, p.ea, p.wa, ld_sa, st);
    if (err != hipSuccess) return err;
    // op(B) column j: storage column j (N); storage row j (T)
    err = prep_operand(g.B, g.ldb, (int)rb, (int)cb, g.tb ? 0 : 1, s, g.batch, p.eb, p.wb, ld_sb, st);
    if (err != hipSuccess) return err;

    // slice operand views: A-T / B-N are K-contiguous
    if (!g.ta && !g.tb) return run_ozaki_gemm<false, true>(p, st);
    if (!g.ta && g.tb) return run_ozaki_gemm<false, false>(p, st);
    if (g.ta && !g.tb) return run_ozaki_gemm<true, true>(p, st);
    return run_ozaki_gemm<true, false>(p, st);
}

}  // namespace ssb
