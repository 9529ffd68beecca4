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
    // error report and slice choice (see oz_bound): per instance the bound
    // (nullptr: not reported) and the slice count used; auto_s: choose the
    // slice count per instance (p.s is then the most allowed)
    float* bound;
    int8_t* sarr;
    float target;
    int auto_s;
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
// s: the most slices (the workspace's plane count); sarr[inst] (when given)
// the instance's own count: planes past it are not written (nor read).
__global__ void __launch_bounds__(256) ozaki_slice(const void* const* X, long ld, int nrows, int ncols,
                                                   const int32_t* e, int exp_by_col, int s, const int8_t* sarr,
                                                   int8_t* w, long ldw)
{
    const int inst = blockIdx.z;
    const int si = sarr != nullptr ? sarr[inst] : s;
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
    for (int p = 0; p < si; ++p) {
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
// per-instance error bound and slice choice of the general path (the bound
// of oz_bound below, over rows and columns of any count): one workgroup per
// instance, one pass over op(A) and op(B) with the exponents of the exponent
// kernels.  Row sums are kept relative to the largest exponent seen so far
// (rescaled by 4^(old - new) when a larger one comes), so one pass suffices.
// ---------------------------------------------------------------------------
constexpr int OZ_SMIN = 3, OZ_NCAND = 9 - OZ_SMIN + 1;

struct OzRowSums {
    int emax = INT32_MIN;          // largest exponent of a nonzero row so far
    float W = 0.f, L2 = 0.f;       // sum w, sum w a2
    float A1 = 0.f, A1sq = 0.f;    // sum w a1, sum w a1^2
    float X1[OZ_NCAND] = {}, X2[OZ_NCAND] = {};  // per candidate s: sum w x, sum w x^2
    __device__ void rescale(int e)
    {
        if (e <= emax) return;
        const float f = emax == INT32_MIN ? 0.f : __builtin_ldexpf(1.f, 2 * max(emax - e, -63));
        W *= f;
        L2 *= f;
        A1 *= f;
        A1sq *= f;
#pragma unroll
        for (int c = 0; c < OZ_NCAND; ++c) {
            X1[c] *= f;
            X2[c] *= f;
        }
        emax = e;
    }
};
struct OzColSums {
    int emax = INT32_MIN;
    float W = 0.f, L2 = 0.f, B1 = 0.f, B1sq = 0.f;
    __device__ void rescale(int e)
    {
        if (e <= emax) return;
        const float f = emax == INT32_MIN ? 0.f : __builtin_ldexpf(1.f, 2 * max(emax - e, -63));
        W *= f;
        L2 *= f;
        B1 *= f;
        B1sq *= f;
        emax = e;
    }
};

__device__ __forceinline__ float oz_cand_c(int s) { return __builtin_ldexpf(1.f / 127.f, 7 * s - 53); }

// block-wide (256 threads) sum and max through LDS
__device__ __forceinline__ float oz_block_sum(float v, float* red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}
__device__ __forceinline__ int oz_block_max(int v, int* red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return max(max(red[0], red[1]), max(red[2], red[3]));
}

__global__ void __launch_bounds__(256) ozaki_stats(const OzParams p)
{
    __shared__ float redf[4];
    __shared__ int redi[4];
    const int inst = blockIdx.x, tid = threadIdx.x;
    const double* A = static_cast<const double*>(p.A[inst]);
    const double* B = static_cast<const double*>(p.B[inst]);
    const int32_t* ea = p.ea + (long)inst * p.m;
    const int32_t* eb = p.eb + (long)inst * p.n;
    const int smax = p.s;
    bool bad = false;
    int Da = 0, Db = 0;  // most digits any value of op(A) / op(B) needs (10: more than nine)
    OzRowSums R;
    for (int i = tid; i < p.m; i += 256) {
        const int e = ea[i];
        if (e == OZ_NONFINITE) {
            bad = true;
            continue;
        }
        float a1 = 0.f, a2 = 0.f, d[9] = {};
        for (int kk = 0; kk < p.k; ++kk) {
            const double x = *gbl(A + (p.ta ? (long)kk + (long)i * p.lda : (long)i + (long)kk * p.lda));
            double v = ldexp(fabs(x), -e);
            const float vf = (float)v;
            a1 += vf;
            a2 = __builtin_fmaf(vf, vf, a2);
            int depth = 0;
#pragma unroll
            for (int q = 0; q < 9; ++q) {  // the slice kernel's digits, magnitudes only
                v *= 128.0;
                const double t = trunc(v);
                v -= t;
                d[q] += (float)t;
                if (t != 0.0) depth = q + 1;
            }
            Da = max(Da, v != 0.0 ? 10 : depth);
        }
        if (a1 == 0.f) continue;  // all-zero row: its products are exact
        R.rescale(e);
        const float w = __builtin_ldexpf(1.f, 2 * max(e - R.emax, -63));
        R.W += w;
        R.L2 += w * a2;
        R.A1 += w * a1;
        R.A1sq += w * a1 * a1;
        float S = 0.f;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            S += d[q];
            if (q + 1 >= OZ_SMIN) {
                const float x = S * (1.f / 128.f) + oz_cand_c(q + 1) * a1;
                R.X1[q + 1 - OZ_SMIN] += w * x;
                R.X2[q + 1 - OZ_SMIN] += w * x * x;
            }
        }
    }
    OzColSums Cs;
    for (int j = tid; j < p.n; j += 256) {
        const int e = eb[j];
        if (e == OZ_NONFINITE) {
            bad = true;
            continue;
        }
        float b1 = 0.f, b2 = 0.f;
        for (int kk = 0; kk < p.k; ++kk) {
            const double x = *gbl(B + (p.tb ? (long)j + (long)kk * p.ldb : (long)kk + (long)j * p.ldb));
            double v = ldexp(fabs(x), -e);
            const float vf = (float)v;
            b1 += vf;
            b2 = __builtin_fmaf(vf, vf, b2);
            int depth = 0;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                v *= 128.0;
                const double t = trunc(v);
                v -= t;
                if (t != 0.0) depth = q + 1;
            }
            Db = max(Db, v != 0.0 ? 10 : depth);
        }
        if (b1 == 0.f) continue;
        Cs.rescale(e);
        const float w = __builtin_ldexpf(1.f, 2 * max(e - Cs.emax, -63));
        Cs.W += w;
        Cs.L2 += w * b2;
        Cs.B1 += w * b1;
        Cs.B1sq += w * b1 * b1;
    }
    bad = oz_block_max(bad ? 1 : 0, redi) != 0;
    Da = oz_block_max(Da, redi);
    Db = oz_block_max(Db, redi);
    const int emax = oz_block_max(R.emax, redi), fmax = oz_block_max(Cs.emax, redi);
    float bound = 0.f;
    int s = smax;
    if (bad) {
        bound = __builtin_huge_valf();
    } else if (emax == INT32_MIN || fmax == INT32_MIN) {
        s = p.auto_s ? OZ_SMIN : smax;  // op(A) or op(B) all zero: exact zero product
    } else {
        R.rescale(emax);
        Cs.rescale(fmax);
        const float Wa = oz_block_sum(R.W, redf), A2 = oz_block_sum(R.L2, redf);
        const float Wb = oz_block_sum(Cs.W, redf), B2 = oz_block_sum(Cs.L2, redf);
        const float B1 = oz_block_sum(Cs.B1, redf), B1sq = oz_block_sum(Cs.B1sq, redf);
        float X1[OZ_NCAND], X2[OZ_NCAND];
#pragma unroll
        for (int c = 0; c < OZ_NCAND; ++c) {
            X1[c] = oz_block_sum(R.X1[c], redf);
            X2[c] = oz_block_sum(R.X2[c], redf);
        }
        const float inv = 1.f / (A2 * B2);
        const float A1 = oz_block_sum(R.A1, redf);
        const float A1sq = oz_block_sum(R.A1sq, redf);
        auto bound_at = [&](int sc) {
            const float c = oz_cand_c(sc);
            const int ci = sc - OZ_SMIN;
            float q;
            if (Da + Db - 1 <= sc) {  // exactly representable: the fold term only
                q = c * c * (A1sq * Wb + 2.f * A1 * B1 + Wa * B1sq);
            } else {
                q = X2[ci] * Wb + 2.f * X1[ci] * (1.f + c) * B1 + Wa * (1.f + c) * (1.f + c) * B1sq;
            }
            return __builtin_ldexpf(__builtin_sqrtf(q * inv), -7 * sc);
        };
        // p.s is the fixed count or the most allowed (at least OZ_SMIN)
        float b = 0.f;
#pragma unroll
        for (int sc = OZ_SMIN; sc <= 9; ++sc)
            if (sc == smax) b = bound_at(sc);
        bound = b;
        if (p.auto_s && b <= p.target) {
#pragma unroll
            for (int sc = 8; sc >= OZ_SMIN; --sc) {
                if (sc >= smax) continue;
                const float bs = bound_at(sc);
                if (!(bs <= p.target)) break;
                s = sc;
                bound = bs;
            }
        }
    }
    if (tid == 0) {
        p.sarr[inst] = (int8_t)s;
        if (p.bound != nullptr) p.bound[inst] = bound;
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
        const int si = p.sarr != nullptr ? p.sarr[inst] : s;  // this instance's slice count
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
            for (int tt = si + 1; tt >= 2; --tt) {
                const int plo = max(1, tt - si), phi = min(si, tt - 1);
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
//
// Statistics for the error bound (oz_bound): l1 += sum of the 16 scaled
// magnitudes |v|, l2 += sum of v^2, dsum[p] += sum of the 16 magnitudes of
// digit p (v_sad_u8 against zero adds the four bytes of a word).
// Taken from the digit words, so that gathering them costs a few byte-dot
// and byte-sum instructions per word: with d_p the p-th digit of a value,
//   l1 = sum (d1 2^-7 + d2 2^-14 + d3 2^-21) + 16 * 2^-21  >= sum |v|
//   l2 = sum (d1 2^-7 + d2 2^-14)^2                      <= sum v^2
// over the lane's 16 values (truncated digits never exceed the value), so the
// bound built on them stays an upper bound.
struct OzStats {
    uint32_t d1 = 0, d2 = 0, d3 = 0;     // digit sums
    uint32_t u11 = 0, u12 = 0, u22 = 0;  // sums of d1 d1, d1 d2, d2 d2
    __device__ float l1() const
    {
        return d1 == 0u ? 0.f : (float)d1 * 0x1p-7f + (float)d2 * 0x1p-14f + (float)d3 * 0x1p-21f + 16.f * 0x1p-21f;
    }
    __device__ float l2() const { return ((float)u11 * 16384.f + (float)u12 * 256.f + (float)u22) * 0x1p-28f; }
};
template <int S, bool STATS = false, bool DIGITS = false>
__device__ __forceinline__ void oz_slice_regs(const double* x, int e, u32x4 (&sl)[S], OzStats* st = nullptr,
                                              uint32_t* dsum = nullptr)
{
    static_assert(S >= 3 && S <= 9, "3 to 9 slices");
    // four values (one 32-bit word of each slice) at a time, so that only
    // their digit fields are live
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t H[4], L[4], N[4], neg = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 4 * j + q;
            const double a = (e == OZ_NONFINITE) ? 0.0 : ldexp(fabs(x[i]), -e);
            neg |= (x[i] < 0.0 ? 0xffu : 0u) << (8 * q);
            const double y1 = ldexp(a, 28);
            H[q] = (uint32_t)y1;
            if constexpr (S > 4) {
                const double y2 = ldexp(y1 - (double)H[q], 28);
                L[q] = (uint32_t)y2;
                if constexpr (S > 8) N[q] = (uint32_t)ldexp(y2 - (double)L[q], 7);
            }
        }
        uint32_t w[9];
        {
            const uint32_t in[4] = {oz_spread7(H[0]), oz_spread7(H[1]), oz_spread7(H[2]), oz_spread7(H[3])};
            uint32_t out[4];
            transpose4x4_bytes(in, out);  // out[c]: byte c of the four values
            w[0] = out[3];
            w[1] = out[2];
            w[2] = out[1];
            w[3] = out[0];
        }
        if constexpr (S > 4) {
            const uint32_t in[4] = {oz_spread7(L[0]), oz_spread7(L[1]), oz_spread7(L[2]), oz_spread7(L[3])};
            uint32_t out[4];
            transpose4x4_bytes(in, out);
            w[4] = out[3];
            w[5] = out[2];
            w[6] = out[1];
            w[7] = out[0];
        }
        if constexpr (S > 8) w[8] = N[0] | (N[1] << 8) | (N[2] << 16) | (N[3] << 24);
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const uint32_t nw = (0x80808080u - w[p]) ^ 0x80808080u;
            sl[p][j] = (neg & nw) | (~neg & w[p]);
            if constexpr (DIGITS) dsum[p] = __builtin_amdgcn_sad_u8(w[p], 0u, dsum[p]);
        }
        if constexpr (STATS) {
            if constexpr (DIGITS) {
                st->d1 = dsum[0];
                st->d2 = dsum[1];
                st->d3 = dsum[2];
            } else {
                st->d1 = __builtin_amdgcn_sad_u8(w[0], 0u, st->d1);
                st->d2 = __builtin_amdgcn_sad_u8(w[1], 0u, st->d2);
                st->d3 = __builtin_amdgcn_sad_u8(w[2], 0u, st->d3);
            }
            st->u11 = __builtin_amdgcn_udot4(w[0], w[0], st->u11, false);
            st->u12 = __builtin_amdgcn_udot4(w[0], w[1], st->u12, false);
            st->u22 = __builtin_amdgcn_udot4(w[1], w[1], st->u22, false);
        }
    }
}

// ---------------------------------------------------------------------------
// a-priori error bound of the Ozaki product, per instance
//
// With v = x * 2^-e (|v| < 1) cut into digits v = sum_{p<=s} v_p 2^-7p + r,
// |r| < 2^-7s, the triangle p + q <= s + 1 computes, for one product a*b,
// sum_p a_p 2^-7p * b_(<= s+1-p)  (b cut after s+1-p digits), so
//     |a*b - computed| <= 2^-7(s+1) sum_p |a_p| + 2^-7s |b|
// (b's remainder after j digits is below 2^-7j; a's after s below 2^-7s).
// The fp64 folds of the orders t = s+1 .. 3 add partial sums below
// 2^-7(t-2) (|a| + |b|) summed over k, so their roundings stay below
// u/127 (a1_i + b1_j); the last fold (t = 2) is the final rounding every
// fp64 result carries and is not part of the bound.  Per element of
// op(A) op(B) (row i, column j; a1_i = sum_k |a'_ik|, b1_j likewise,
// S_i(s) = sum_k sum_{p<=s} |a_p|):
//     |dP_ij| <= 2^(e_i+f_j) 2^-7s (x_i + y_j),
//     x_i = S_i(s)/128 + c_s a1_i,  y_j = (1 + c_s) b1_j,  c_s = 2^(7s-53)/127.
// Summed over (i, j) the square separates into row and column sums:
//     ||dP||_F^2 <= 2^-14s (X2 Wb + 2 X1 (1+c_s) B1 + Wa (1+c_s)^2 B1sq) * 4^(emax+fmax)
// with weights w_i = 4^(e_i-emax) (0 for an all-zero row), Wa = sum w_i,
// X1 = sum w_i x_i, X2 = sum w_i x_i^2, and wb_j, Wb, B1 = sum wb_j b1_j,
// B1sq = sum wb_j b1_j^2 for the columns; ||op(A)||_F^2 = 4^emax sum w_i a2_i
// (a2_i = sum_k a'_ik^2), likewise op(B).  The normwise bound
//     ||dP||_F / (||op(A)||_F ||op(B)||_F)
// is evaluated in fp32 (an estimate of a bound: 2^-24 relative is ample).
// Exactly representable operands: when every value of op(A) needs at most
// Da digits and every value of op(B) at most Db, and Da + Db - 1 <= s, no
// digit is cut and no product of the triangle is dropped, so only the fold
// term remains (x_i = c_s a1_i, y_j = c_s b1_j).
// Auto mode takes the smallest s (<= p.s) whose bound is at most the
// target; an instance that misses it at p.s keeps p.s and reports its bound.
// Rows or columns with inf / nan: bound +inf (their outputs are NaN).
// ---------------------------------------------------------------------------
constexpr float OZ_TARGET = 2.7755575615628914e-17f;  // 2^-55, a quarter unit roundoff of fp64

// Sum over the 32 row (column) owners; lanes l and l ^ 32 hold the same values.
__device__ __forceinline__ float oz_sum32(float v)
{
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ int oz_max32(int v)
{
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// Bound at S slices for the fused one-wave form: lane l owns row l % 32 of
// op(A) and column l % 32 of op(B) over the k half l / 32 (the statistics of
// both halves are joined first).  The fused kernel's time is set by its
// loads and its slicing, not by the slice count, so its automatic choice is
// always the most slices and only the bound at S is evaluated here; the
// general path chooses per instance (ozaki_stats).
template <int S>
__device__ __forceinline__ float oz_bound(int ea, int eb, OzStats sa, OzStats sb, const uint32_t (&dsum)[S])
{
    if (__any(ea == OZ_NONFINITE || eb == OZ_NONFINITE)) return __builtin_huge_valf();
    // join the two k halves
    float a1 = sa.l1(), a2 = sa.l2(), b1 = sb.l1(), b2 = sb.l2();
    a1 += __shfl_xor(a1, 32);
    a2 += __shfl_xor(a2, 32);
    b1 += __shfl_xor(b1, 32);
    b2 += __shfl_xor(b2, 32);
    uint32_t Sa = 0;  // digit sum S_i(S)
#pragma unroll
    for (int p = 0; p < S; ++p) Sa += dsum[p];
    Sa += (uint32_t)__shfl_xor((int)Sa, 32);
    const bool rowz = a1 == 0.f, colz = b1 == 0.f;
    const int emax = oz_max32(rowz ? INT32_MIN : ea), fmax = oz_max32(colz ? INT32_MIN : eb);
    if (emax == INT32_MIN || fmax == INT32_MIN) return 0.f;  // op(A) or op(B) all zero: exact zero product
    const float wa = rowz ? 0.f : __builtin_ldexpf(1.f, 2 * max(ea - emax, -63));
    const float wb = colz ? 0.f : __builtin_ldexpf(1.f, 2 * max(eb - fmax, -63));
    const float Wa = oz_sum32(wa), A2 = oz_sum32(wa * a2);
    const float Wb = oz_sum32(wb), B2 = oz_sum32(wb * b2);
    const float B1 = oz_sum32(wb * b1), B1sq = oz_sum32(wb * b1 * b1);
    const float c = __builtin_ldexpf(1.f / 127.f, 7 * S - 53);
    const float x = (float)Sa * (1.f / 128.f) + c * a1;
    const float cy = 1.f + c;
    const float X1 = oz_sum32(wa * x), X2 = oz_sum32(wa * x * x);
    const float q = X2 * Wb + 2.f * X1 * cy * B1 + Wa * cy * cy * B1sq;
    return __builtin_ldexpf(__builtin_sqrtf(q / (A2 * B2)), -7 * S);
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
// The lane's offsets go through an empty asm per instance: left visible, the
// compiler hoists all 16 of them out of the instance loop and keeps them
// live in registers across it.
__device__ __forceinline__ uint32_t oz_opaque(uint32_t x)
{
    __asm__ volatile("" : "+v"(x));
    return x;
}
template <bool FULL>
__device__ __forceinline__ void oz_small_load(const void* X, const OzLane& L, double (&x)[16])
{
    const double* P = wave_uniform(static_cast<const double*>(X));
    const uint32_t base = oz_opaque(L.base), step = oz_opaque(L.step);
    if constexpr (FULL) {
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = *gbl(P + (base + (uint32_t)q * step));
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t qq = (uint32_t)q < L.kvalid ? (uint32_t)q : 0u;  // stays inside the matrix
            const double v = *gbl(P + (base + qq * step));
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
        OzStats sta, stb;
        uint32_t dsum[S];
#pragma unroll
        for (int q = 0; q < S; ++q) dsum[q] = 0u;
        const int ea = oz_exp16(xa);
        oz_slice_regs<S, true, true>(xa, ea, sa, &sta, dsum);
        const int eb = oz_exp16(xb);
        oz_slice_regs<S, true, false>(xb, eb, sb, &stb);

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
        // the instance's error bound, from the statistics gathered while
        // slicing (lanes beyond the instance hold zeros); computed after the
        // stores so that its reductions do not add to the product loop's
        // registers
        if (p.bound != nullptr) {
            const float bound = oz_bound<S>(ea, eb, sta, stb, dsum);
            if (lane == 0) {
                p.bound[inst] = bound;
                p.sarr[inst] = (int8_t)S;
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

// Exponents of one operand's rows (op(A)) or columns (op(B)).  Storage is
// nrows x ncols (column major); exponents are per storage column when
// exp_by_col.
static hipError_t exp_operand(const void* const* X, long ld, int nrows, int ncols, int exp_by_col, int batch,
                              int32_t* e, hipStream_t st)
{
    if (exp_by_col) {
        hipLaunchKernelGGL(ozaki_exp_cols, dim3((ncols + 3) / 4, batch), dim3(256), 0, st, X, ld, nrows, ncols, e);
    } else {
        hipLaunchKernelGGL(ozaki_exp_rows, dim3((nrows + 63) / 64, batch), dim3(256), 0, st, X, ld, nrows, ncols, e);
    }
    return hipGetLastError();
}

static hipError_t slice_operand(const void* const* X, long ld, int nrows, int ncols, int exp_by_col, int s,
                                const int8_t* sarr, int batch, const int32_t* e, int8_t* w, long ldw,
                                hipStream_t st)
{
    hipLaunchKernelGGL(ozaki_slice, dim3((unsigned)((ldw / 4 + 255) / 256), ncols, batch), dim3(256), 0, st, X, ld,
                       nrows, ncols, e, exp_by_col, s, sarr, w, ldw);
    return hipGetLastError();
}

hipError_t ozaki_run(const GemmProblem& g, int slices, WsGet ws_get, hipStream_t st, OzakiReportLoc* rep)
{
    const int m = g.m, n = g.n, k = g.k;
    const bool auto_s = slices == 0;
    const int s = auto_s ? 9 : slices;  // the most slices an instance may take
    const OzakiLayout L = ozaki_layout(g, s);
    hipError_t err = hipSuccess;
    char* ws = static_cast<char*>(ws_get(L.total, &err));
    if (ws == nullptr) return err == hipSuccess ? hipErrorOutOfMemory : err;

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
    p.sc.af = g.af;
    p.sc.bf = g.bf;
    p.sc.ad = g.ad;
    p.sc.bd = g.bd;
    p.sc.ai = g.ai;
    p.sc.bi = g.bi;
    p.beta0 = g.beta0 ? 1 : 0;
    p.bound = reinterpret_cast<float*>(ws + L.off_bound);
    p.sarr = reinterpret_cast<int8_t*>(ws + L.off_s);
    p.target = OZ_TARGET;
    p.auto_s = auto_s ? 1 : 0;
    rep->bound = p.bound;
    rep->slices = p.sarr;
    rep->batch = g.batch;

    if (ozaki_fused_shape(g)) return ozaki_small_dispatch(p, st);  // slices in registers: no more workspace

    const long ra = g.ta ? k : m, ca = g.ta ? m : k;  // storage of A
    const long rb = g.tb ? n : k, cb = g.tb ? k : n;  // storage of B
    p.wa = reinterpret_cast<int8_t*>(ws + L.off_a);
    p.wb = reinterpret_cast<int8_t*>(ws + L.off_b);
    p.ea = reinterpret_cast<int32_t*>(ws + L.off_e);
    p.eb = p.ea + (size_t)g.batch * m;
    p.ld_sa = L.ld_sa;
    p.cols_sa = ca;
    p.ld_sb = L.ld_sb;
    p.cols_sb = cb;

    // op(A) row i: storage row i (N) -> reduce along the strided dim; storage column i (T)
    err = exp_operand(g.A, g.lda, (int)ra, (int)ca, g.ta ? 1 : 0, g.batch, p.ea, st);
    if (err != hipSuccess) return err;
    // op(B) column j: storage column j (N); storage row j (T)
    err = exp_operand(g.B, g.ldb, (int)rb, (int)cb, g.tb ? 0 : 1, g.batch, p.eb, st);
    if (err != hipSuccess) return err;
    // error bound and slice count per instance
    hipLaunchKernelGGL(ozaki_stats, dim3((unsigned)g.batch), dim3(256), 0, st, p);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    err = slice_operand(g.A, g.lda, (int)ra, (int)ca, g.ta ? 1 : 0, s, p.sarr, g.batch, p.ea, p.wa, L.ld_sa, st);
    if (err != hipSuccess) return err;
    err = slice_operand(g.B, g.ldb, (int)rb, (int)cb, g.tb ? 0 : 1, s, p.sarr, g.batch, p.eb, p.wb, L.ld_sb, st);
    if (err != hipSuccess) return err;

    // slice operand views: A-T / B-N are K-contiguous
    if (!g.ta && !g.tb) return run_ozaki_gemm<false, true>(p, st);
    if (!g.ta && g.tb) return run_ozaki_gemm<false, false>(p, st);
    if (g.ta && !g.tb) return run_ozaki_gemm<true, true>(p, st);
    return run_ozaki_gemm<true, false>(p, st);
}

}  // namespace ssb
