// ozaki.hip -- fp64 batched GEMM emulated on the int8 MFMA units (Ozaki
// scheme with integer slices: Ozaki, Ogita, Oishi, Rump, Numer. Algorithms
// 59 (2012); int8 form after Ootomo, Ozaki, Yokota, IJHPCA 2024).
//
// For every instance:
//   1. e_i = exponent with max_k |op(A)(i,k)| < 2^e_i  (row i of op(A)),
//      f_j likewise for column j of op(B)
//   2. A'(i,k) = op(A)(i,k) * 2^-e_i in (-1, 1) is cut into s signed 7-bit
//      slices:  r = A' ; repeat s times { r *= 128; a_p = trunc(r); r -= a_p }
//      so A' = sum_p a_p 2^(-7p) + O(2^(-7s)), |a_p| <= 127; same for B.
//      Every step is exact in fp64.
//   3. for t = s+1 down to 2:  D_t = sum_{p+q=t} a_p b_q  -- exact int32
//      products on the int8 MFMA -- and D += D_t * 2^(-7t) in fp64
//      (smallest terms first)
//   4. C = fma(D * 2^(e_i+f_j), alpha, fl(beta * C))  (the reference's
//      epilogue rounding order)
// The triangle p + q <= s + 1 keeps s(s+1)/2 of the s^2 slice products; the
// dropped ones are O(2^(-7(s+1))) relative to the row/column maxima.
// int32 exactness: one D_t sums at most s*kc*127^2 < 2^31 for a K chunk kc
// of at most 8192; longer K is folded chunk by chunk.
// Rows or columns holding inf/nan get NaN in every output they touch.
//
// Two forms compute the same result bit for bit:
//   * instances up to 32 x 32 x 32: ozaki_small, one wave per instance,
//     slices cut in registers;
//   * larger ones: ozaki_prep (exponents, slices into the workspace, row
//     statistics; one launch per operand), ozaki_choose (error bound and
//     slice count per instance), ozaki_mma (slice products and fold).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>

#include <algorithm>

#include "gemm_dma.hpp"
#include "api_internal.hpp"
#include "launch.hpp"
#include "occupancy.hpp"

namespace ssb {

struct OzRow;

struct OzParams {
    const void* const* A;  // caller's fp64 operands
    const void* const* B;
    void* const* C;
    long lda, ldb, ldc;
    int m, n, k, batch, s;
    int ta, tb;
    // workspace of the general path (see ozaki_prep)
    int8_t* wa;      // slices of op(A): [inst][kb][plane][rows_pad_a][OZ_KB]
    int8_t* wb;      // slices of op(B): [inst][kb][plane][rows_pad_b][OZ_KB]
                     // (kb: k-blocks of OZ_KB, an even count)
    int32_t* ea;     // [inst][m] exponents (INT_MIN marks a non-finite row)
    int32_t* eb;     // [inst][n]
    OzRow* ra;       // [inst][m] row statistics of op(A)
    OzRow* rb;       // [inst][n] column statistics of op(B)
    long rows_pad_a, rows_pad_b;  // m, n rounded up to OZ_ROWS
    long inst_a, inst_b;          // slice bytes per instance
    int nkb, planes;              // k-blocks, slice planes per k-block (= s)
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

typedef int32_t i32x16 __attribute__((ext_vector_type(16)));

constexpr int OZ_NONFINITE = INT32_MIN;
constexpr int OZ_KCHUNK = 8192;

__device__ __forceinline__ int exp_of_max(double mx, bool nonfinite)
{
    if (nonfinite) return OZ_NONFINITE;
    if (mx == 0.0) return 0;
    int e;
    (void)frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1)  ->  mx < 2^e
    return e;
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

// The 7-bit digits of a 28-bit value, one per byte: byte 3 the most
// significant digit, byte 0 the least.
__device__ __forceinline__ uint32_t oz_spread7(uint32_t h)
{
    return (h & 0x7fu) | ((h << 1) & 0x7f00u) | ((h << 2) & 0x7f0000u) | ((h << 3) & 0x7f000000u);
}

// ---------------------------------------------------------------------------
// General path (any dimension above 32).
//
// Workspace per instance: the exponents of op(A)'s rows and op(B)'s columns,
// their statistics (OzRow), and the slices, k-block major: plane p of
// k-block kb of an operand is rows_pad x OZ_KB bytes, row r's digits of
// k = OZ_KB kb .. OZ_KB kb + OZ_KB - 1 contiguous:
//     slices[inst][kb][p][r][kk]
// ("rows" are op(A)'s rows, or op(B)'s columns).  Rows past the operand and
// k past K (up to a multiple of 64) hold zero digits, so the product
// kernel's k-blocks are whole and its loads need no guards.
// ---------------------------------------------------------------------------
constexpr int OZ_ROWS = 64;  // rows per prep workgroup and per product tile side
constexpr int OZ_KB = 32;    // k per block (bytes of one slice row; one MFMA k-step)

// Per row of op(A) (column of op(B)), over k: the scaled values v = x 2^-e
struct OzRow {
    float a1, a2;  // sum |v|, sum v^2 (fp32)
    float d[9];    // sum of digit p + 1 (magnitudes)
    float depth;   // most digits a value of the row needs (10: more than nine)
};
static_assert(sizeof(OzRow) == OZAKI_ROW_BYTES, "row statistics layout");

struct OzPrep {
    const void* const* X;
    long ld;
    int rows, k;      // rows of op(X) (op(B): columns), K
    int nrb;          // row blocks of OZ_ROWS per instance
    int planes;       // slice planes written per k-block
    long rows_pad;    // nrb * OZ_ROWS
    long inst_bytes;  // slice bytes per instance
    int8_t* w;
    int32_t* e;       // [inst][rows]
    OzRow* st;        // [inst][rows]
};

// 16 values of row r, k = k0 .. k0 + 15 (zeros past K).  KC: element (r, k)
// at r * ld + k, else at r + k * ld.  Indices are clamped into the matrix,
// so every load is in bounds.
template <bool KC>
__device__ __forceinline__ void oz_load16(const double* X, long ld, int r, int k0, int K, double (&x)[16])
{
    if (KC) {
        const double* row = X + (long)r * ld;
        if (k0 + 16 <= K && aligned16(row + k0)) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const f64x2 v = *gbl(reinterpret_cast<const f64x2*>(row + k0) + i);
                x[2 * i] = v[0];
                x[2 * i + 1] = v[1];
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const double v = *gbl(row + min(k0 + i, K - 1));
            x[i] = k0 + i < K ? v : 0.0;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const double v = *gbl(X + r + (long)min(k0 + i, K - 1) * ld);
            x[i] = k0 + i < K ? v : 0.0;
        }
    }
}

// The nine digits of 16 values of one row (row exponent e) into sl[p]
// (16 bytes: digit p + 1 of the 16 values, signed), as the sequence
// "v *= 128; d = trunc(v); v -= d" gives them; done in integers as in the
// fused kernel (oz_slice_regs).  Row statistics: a1, a2 (fp32 sums of |v|,
// v^2), dsum[p] (digit magnitudes), rem (a value has digits past the ninth).
__device__ __forceinline__ void oz_slice16(const double (&x)[16], int e, u32x4 (&sl)[9], float& a1, float& a2,
                                           uint32_t (&dsum)[9], bool& rem)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t H[4], L[4], N[4], neg = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 4 * j + q;
            const double a = (e == OZ_NONFINITE) ? 0.0 : ldexp(fabs(x[i]), -e);
            const float af = (float)a;
            a1 += af;
            a2 = __builtin_fmaf(af, af, a2);
            neg |= (x[i] < 0.0 ? 0xffu : 0u) << (8 * q);
            // remainders by v_fract_f64 (exact for these non-negative
            // values below 2^28)
            const double y1 = ldexp(a, 28);
            H[q] = (uint32_t)y1;
            const double y2 = ldexp(__builtin_amdgcn_fract(y1), 28);
            L[q] = (uint32_t)y2;
            const double y3 = ldexp(__builtin_amdgcn_fract(y2), 7);
            N[q] = (uint32_t)y3;
            rem |= __builtin_amdgcn_fract(y3) != 0.0;
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
        {
            const uint32_t in[4] = {oz_spread7(L[0]), oz_spread7(L[1]), oz_spread7(L[2]), oz_spread7(L[3])};
            uint32_t out[4];
            transpose4x4_bytes(in, out);
            w[4] = out[3];
            w[5] = out[2];
            w[6] = out[1];
            w[7] = out[0];
        }
        w[8] = N[0] | (N[1] << 8) | (N[2] << 16) | (N[3] << 24);
#pragma unroll
        for (int p = 0; p < 9; ++p) {
            const uint32_t nw = (0x80808080u - w[p]) ^ 0x80808080u;
            sl[p][j] = (neg & nw) | (~neg & w[p]);
            dsum[p] = __builtin_amdgcn_sad_u8(w[p], 0u, dsum[p]);
        }
    }
}

// One workgroup per (instance, block of 64 rows): thread t takes row t / 4
// of the block, k = 16 (t % 4) .. + 15 of every 64 k.  Pass 1: the row
// maximum -> exponent; pass 2: digits, written as `planes` slice planes,
// and the row statistics.  Per 64 k and plane the workgroup stores two
// contiguous 2 KiB k-blocks.  NK > 0 (K <= 64 NK): pass 1 keeps the thread's
// values in registers for pass 2, so the operand is read once; NK == 0:
// pass 2 reads it again.
template <bool KC, int NK>
__global__ void __launch_bounds__(256) ozaki_prep(const OzPrep q)
{
    const int inst = blockIdx.x / q.nrb, rblk = blockIdx.x % q.nrb;
    const int tid = threadIdx.x, h = tid & 3;
    const int r = rblk * OZ_ROWS + (tid >> 2);
    const bool row_ok = r < q.rows;
    const int rr = row_ok ? r : q.rows - 1;
    const double* X = static_cast<const double*>(q.X[inst]);
    const int nk64 = (q.k + 63) / 64;

    // Row exponent as exp_of_max gives it, from the largest biased exponent
    // field E of the row (the high words of |x| compare as the values do):
    // e = E - 1022 for a normal maximum, OZ_NONFINITE for E = 0x7ff (inf /
    // nan); a zero or subnormal maximum (E = 0) goes through frexp.
    uint32_t hmax = 0u;
    auto hi_abs = [](double v) { return (uint32_t)(__builtin_bit_cast(uint64_t, v) >> 32) & 0x7fffffffu; };
    double xs[NK > 0 ? NK : 1][16];
    if constexpr (NK > 0) {
#pragma unroll
        for (int kb = 0; kb < NK; ++kb) {
            if (kb < nk64) oz_load16<KC>(X, q.ld, rr, kb * 64 + 16 * h, q.k, xs[kb]);
        }
#pragma unroll
        for (int kb = 0; kb < NK; ++kb) {
            if (kb < nk64) {
#pragma unroll
                for (int i = 0; i < 16; ++i) hmax = max(hmax, hi_abs(xs[kb][i]));
            }
        }
    } else {
        for (int kb = 0; kb < nk64; ++kb) {
            double x[16];
            oz_load16<KC>(X, q.ld, rr, kb * 64 + 16 * h, q.k, x);
#pragma unroll
            for (int i = 0; i < 16; ++i) hmax = max(hmax, hi_abs(x[i]));
        }
    }
#pragma unroll
    for (int off = 1; off < 4; off <<= 1) hmax = max(hmax, (uint32_t)__shfl_xor((int)hmax, off));
    const int E = (int)(hmax >> 20);
    int e = E == 0x7ff ? OZ_NONFINITE : E - 1022;
    if (E == 0) {
        double mx = 0.0;
        if constexpr (NK > 0) {
#pragma unroll
            for (int kb = 0; kb < NK; ++kb)
                if (kb < nk64) {
#pragma unroll
                    for (int i = 0; i < 16; ++i) mx = fmax(mx, fabs(xs[kb][i]));
                }
        } else {
            for (int kb = 0; kb < nk64; ++kb) {
                double x[16];
                oz_load16<KC>(X, q.ld, rr, kb * 64 + 16 * h, q.k, x);
#pragma unroll
                for (int i = 0; i < 16; ++i) mx = fmax(mx, fabs(x[i]));
            }
        }
        // the four threads of a row take this branch together (E is the
        // row's)
#pragma unroll
        for (int off = 1; off < 4; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off));
        e = exp_of_max(mx, false);
    }
    if (!row_ok) e = OZ_NONFINITE;  // padding rows: zero digits

    float a1 = 0.f, a2 = 0.f;
    uint32_t dsum[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) dsum[p] = 0u;
    bool rem = false;
    const long plane = q.rows_pad * OZ_KB;
    // k-block 2 kb + h / 2 of 32, bytes 16 (h % 2) .. of the row
    int8_t* out = q.w + (long)inst * q.inst_bytes + (h >> 1) * q.planes * plane +
                  ((long)rblk * OZ_ROWS + (tid >> 2)) * OZ_KB + (h & 1) * 16;
    auto slice_out = [&](const double(&x)[16], int kb) {
        u32x4 sl[9];
        oz_slice16(x, e, sl, a1, a2, dsum, rem);
        int8_t* o = out + (long)(2 * kb) * q.planes * plane;
#pragma unroll
        for (int p = 0; p < 9; ++p)
            if (p < q.planes) *gbl(reinterpret_cast<u32x4*>(o + p * plane)) = sl[p];
    };
    if constexpr (NK > 0) {
#pragma unroll
        for (int kb = 0; kb < NK; ++kb)
            if (kb < nk64) slice_out(xs[kb], kb);
    } else {
        for (int kb = 0; kb < nk64; ++kb) {
            double x[16];
            oz_load16<KC>(X, q.ld, rr, kb * 64 + 16 * h, q.k, x);
            slice_out(x, kb);
        }
    }
#pragma unroll
    for (int off = 1; off < 4; off <<= 1) {
        a1 += __shfl_xor(a1, off);
        a2 += __shfl_xor(a2, off);
        rem |= __shfl_xor((int)rem, off) != 0;
#pragma unroll
        for (int p = 0; p < 9; ++p) dsum[p] += (uint32_t)__shfl_xor((int)dsum[p], off);
    }
    if (h == 0 && row_ok) {
        const long idx = (long)inst * q.rows + r;
        q.e[idx] = e;
        int depth = 0;
#pragma unroll
        for (int p = 0; p < 9; ++p)
            if (dsum[p] != 0u) depth = p + 1;
        OzRow st;
        st.a1 = a1;
        st.a2 = a2;
#pragma unroll
        for (int p = 0; p < 9; ++p) st.d[p] = (float)dsum[p];
        st.depth = (float)(rem ? 10 : depth);
        float4* dst = reinterpret_cast<float4*>(q.st + idx);
        const float4* src = reinterpret_cast<const float4*>(&st);
        dst[0] = src[0];
        dst[1] = src[1];
        dst[2] = src[2];
    }
}

// ---------------------------------------------------------------------------
// per-instance error bound and slice choice of the general path (the bound
// of oz_bound below, over rows and columns of any count): one workgroup per
// instance over the row statistics of ozaki_prep.  Row sums are kept
// relative to the largest exponent seen so far (rescaled by 4^(old - new)
// when a larger one comes), so one pass suffices.
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

__global__ void __launch_bounds__(256) ozaki_choose(const OzParams p)
{
    __shared__ float redf[4];
    __shared__ int redi[4];
    const int inst = blockIdx.x, tid = threadIdx.x;
    const int32_t* ea = p.ea + (long)inst * p.m;
    const int32_t* eb = p.eb + (long)inst * p.n;
    const OzRow* ra = p.ra + (long)inst * p.m;
    const OzRow* rb = p.rb + (long)inst * p.n;
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
        const OzRow r = ra[i];
        Da = max(Da, (int)r.depth);
        if (r.a1 == 0.f) continue;  // all-zero row: its products are exact
        R.rescale(e);
        const float w = __builtin_ldexpf(1.f, 2 * max(e - R.emax, -63));
        R.W += w;
        R.L2 += w * r.a2;
        R.A1 += w * r.a1;
        R.A1sq += w * r.a1 * r.a1;
        float S = 0.f;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            S += r.d[q];
            if (q + 1 >= OZ_SMIN) {
                const float x = S * (1.f / 128.f) + oz_cand_c(q + 1) * r.a1;
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
        const OzRow r = rb[j];
        Db = max(Db, (int)r.depth);
        if (r.a1 == 0.f) continue;
        Cs.rescale(e);
        const float w = __builtin_ldexpf(1.f, 2 * max(e - Cs.emax, -63));
        Cs.W += w;
        Cs.L2 += w * r.a2;
        Cs.B1 += w * r.a1;
        Cs.B1sq += w * r.a1 * r.a1;
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
// Slice products with fp64 fold (general path): one BM x 64 tile of C at a
// time per workgroup, persistent over the tiles; BM / 32 x 2 waves of
// 32 x 32, each on v_mfma_i32_32x32x32_i8 (BM = 64: four waves, one per
// SIMD; BM = 128: eight, two per SIMD -- fewer slice bytes per product).  A
// step is one k-block of 32: every plane of the tile's A rows and B columns
// moves global -> LDS with global_load_lds_dwordx4 into one of the stages.
// The loads run ahead through the workgroup's whole sequence of tiles, so a
// tile's first k-blocks arrive while the previous tile finishes; the tile's
// row and column exponents ride along with its first step (into LDS, so
// that the epilogue never waits on the vector memory counter behind later
// loads).  Every slice pair (p, q) of the triangle p + q <= s + 1 is
// multiplied from LDS into the int32 accumulator of its order t = p + q
// (s accumulators of 16 registers).  After each K chunk of OZ_KCHUNK the
// orders are folded into fp64, smallest first (D = fma(D_t, 2^-7t, D) for
// t = s + 1 .. 2); then the epilogue C = fma(D * 2^(e_i + f_j), alpha,
// fl(beta * C)).
// Every step loads all S planes (an instance with fewer slices multiplies
// only its own), so the steps are equal and the waits fixed counts.
// LDS rows are 32 bytes (two 16-byte chunks); chunk c of row r sits at chunk
// c ^ ((r >> 3) & 1), so a fragment read (rows r0 .. r0 + 15 of one chunk
// column) covers 16 distinct bank groups.
// ---------------------------------------------------------------------------
template <int S, int BM>
struct OzMma {
    static constexpr int WM = BM / 32, NW = 2 * WM;  // waves along M, waves
    static constexpr int PLANE_A = BM * OZ_KB;       // one plane of the tile's A rows, bytes
    static constexpr int PLANE_B = OZ_ROWS * OZ_KB;  // of its B columns
    static constexpr int STAGE = S * (PLANE_A + PLANE_B);
    static constexpr int NB = STAGE / 1024;          // 1 KiB load blocks per step
    static constexpr int EXP_TILE = (BM + OZ_ROWS) * 4;
    static constexpr int NST_FIT = (160 * 1024 - 4 * EXP_TILE) / STAGE;
    static constexpr int NST = NST_FIT < 4 ? NST_FIT : 4;  // stages
    static constexpr int AHEAD = NST - 1;                  // steps loaded ahead
    static constexpr int DPW = NB / NW;                    // loads per wave per step (at least)
    static constexpr int EXP = NST * STAGE;                // exponents: 4 tiles x (BM rows + 64 columns) int32
    static constexpr int LDS_BYTES = EXP + 4 * EXP_TILE;
    static_assert(NST >= 2, "at least two stages");
    static_assert(LDS_BYTES <= 160 * 1024, "stages fit the LDS");
};

__device__ __forceinline__ int oz_chunk(int row, int c) { return row * OZ_KB + ((c ^ ((row >> 3) & 1)) << 4); }

// The workgroup's i-th tile (tiles blockIdx.x, blockIdx.x + gridDim.x, ...;
// xcd_remap gives one XCD neighbouring tiles).  The host keeps the tile
// count below 2^31, so the index arithmetic is 32-bit.
struct OzTile {
    int inst, i0, j0;
};
template <int BM>
__device__ __forceinline__ OzTile oz_tile(const OzParams& p, int i)
{
    const uint32_t t = blockIdx.x + (uint32_t)i * gridDim.x;
    const uint32_t total = (uint32_t)p.total;
    const uint32_t q = total >> 3, r = total & 7, x = t & 7;
    const uint32_t w = x * q + (x < r ? x : r) + (t >> 3);  // xcd_remap
    const uint32_t tpi = (uint32_t)(p.tiles_m * p.tiles_n);
    const uint32_t inst = w / tpi, rem = w - inst * tpi;
    const uint32_t tm = rem % (uint32_t)p.tiles_m;
    return OzTile{(int)inst, (int)tm * BM, (int)((rem - tm) / (uint32_t)p.tiles_m) * OZ_ROWS};
}

// A byte of a read-only device array through the scalar cache (a vector
// load would be ordered behind the outstanding LDS loads on the vector
// memory counter).  The dword holding it is read; the wait is inside.
__device__ __forceinline__ int oz_scalar_byte(const int8_t* base, int idx)
{
    const int8_t* wp = base + (idx & ~3);
    uint32_t w;
    __asm__ volatile("s_load_dword %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(w) : "s"(wp) : "memory");
    return (int)(int8_t)(w >> (8 * (idx & 3)));
}

// Where the loads stand in the workgroup's sequence of steps (step g: tile
// g / nkb, k-block g % nkb), kept incrementally (all wave-uniform).
struct OzStream {
    long g;     // next step to load
    int kb, i;  // its k-block and tile
    const char* a;  // slices of that tile's A rows at k-block kb
    const char* b;  // and of its B columns
    int inst, i0, j0;
};
template <int BM>
__device__ __forceinline__ void oz_stream_tile(const OzParams& p, OzStream& L)
{
    const OzTile T = oz_tile<BM>(p, L.i);
    L.inst = T.inst;
    L.i0 = T.i0;
    L.j0 = T.j0;
    L.a = reinterpret_cast<const char*>(p.wa) + T.inst * p.inst_a + (long)T.i0 * OZ_KB;
    L.b = reinterpret_cast<const char*>(p.wb) + T.inst * p.inst_b + (long)T.j0 * OZ_KB;
}

// Loads of step L.g into stage g % NST, then advance.  1 KiB block blk of
// the stage (blk = wave + NW j): A's planes first (BM / 32 blocks each),
// then B's (two each); a block is 32 rows of one plane, lane l fills row
// l / 2 of it, LDS chunk l % 2, with the source chunk that belongs there.
// A tile's first step also brings its exponents (A rows: waves 0 ..
// BM / 64 - 1, B columns: wave BM / 64).
template <int S, int BM>
__device__ __forceinline__ void oz_issue(const OzParams& p, OzStream& L, long gend, char* smem, int wave, int lane)
{
    using M = OzMma<S, BM>;
    constexpr int AB = S * (BM / 32);  // A's blocks
    const long plane_a = p.rows_pad_a * OZ_KB, plane_b = p.rows_pad_b * OZ_KB;
    char* st = smem + (int)(L.g % M::NST) * M::STAGE;
    const int rl = lane >> 1, cl = lane & 1;
#pragma unroll
    for (int j = 0; j < (M::NB + M::NW - 1) / M::NW; ++j) {
        const int blk = wave + M::NW * j;
        if (blk < M::NB) {
            const char* from;
            if (blk < AB) {
                const int pl = blk / (BM / 32), row = (blk % (BM / 32)) * 32 + rl;
                from = L.a + pl * plane_a + oz_chunk(row, cl);
            } else {
                const int b2 = blk - AB, pl = b2 >> 1, row = (b2 & 1) * 32 + rl;
                from = L.b + pl * plane_b + oz_chunk(row, cl);
            }
            __builtin_amdgcn_global_load_lds(from, (lds_void_t*)(st + blk * 1024), 16, 0, 0);
        }
    }
    if (L.kb == 0 && wave <= BM / 64) {
        // row (waves 0 .. BM / 64 - 1) or column (wave BM / 64) exponents of
        // the tile; lanes past the operand read its last one
        const int32_t* e = wave < BM / 64 ? p.ea + (long)L.inst * p.m + min(L.i0 + wave * 64 + lane, p.m - 1)
                                          : p.eb + (long)L.inst * p.n + min(L.j0 + lane, p.n - 1);
        __builtin_amdgcn_global_load_lds(e, (lds_void_t*)(smem + M::EXP + (L.i & 3) * M::EXP_TILE + wave * 256), 4,
                                         0, 0);
    }
    ++L.g;
    L.a += S * plane_a;
    L.b += S * plane_b;
    if (++L.kb == p.nkb) {
        L.kb = 0;
        ++L.i;
        if (L.g < gend) oz_stream_tile<BM>(p, L);
    }
}

// One tile: steps g .. g + nkb - 1 of the workgroup's sequence (gend steps
// in all).  FULL: the instance takes all S slices (no per-pair test).
// CHUNKS: K spans several chunks (the fp64 sums then stay live across the
// k-loop; with one chunk they only exist after it).
template <int S, int BM, bool FULL, bool CHUNKS>
__device__ __forceinline__ void oz_mma_tile(const OzParams& p, const OzTile& T, int i, int si, char* smem, long& g,
                                            OzStream& L, long gend)
{
    using M = OzMma<S, BM>;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % M::WM, wn = wave / M::WM;
    const int fr_a = wm * 32 + (lane & 31), fr_b = wn * 32 + (lane & 31), kh = lane >> 5;
    const int nkb = p.nkb;

    double D[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) D[v] = 0.0;
    // K chunks: the accumulators are only updated by the MFMAs inside the
    // k-block loop (a fold inside it would make the compiler move all of
    // them between accumulation and ordinary registers every k-block)
    for (int c0 = 0; c0 < nkb; c0 += OZ_KCHUNK / OZ_KB) {
        const int c1 = CHUNKS ? min(nkb, c0 + OZ_KCHUNK / OZ_KB) : nkb;
        i32x16 acc[S];
#pragma unroll
        for (int t = 0; t < S; ++t) acc[t] = i32x16{};
        for (int kb = c0; kb < c1; ++kb, ++g) {
            // step g's loads are complete once at most the later steps'
            // remain in flight (loads and stores retire in order; waiting
            // for more than needed is safe)
            if (g + M::AHEAD - 1 < gend) wait_vm<(M::AHEAD - 1) * M::DPW>();
            else wait_vm<0>();
            lds_barrier();
            if (L.g < gend) oz_issue<S, BM>(p, L, gend, smem, wave, lane);  // step g + AHEAD
            const char* la = smem + (int)(g % M::NST) * M::STAGE;
            const char* lb = la + S * M::PLANE_A;
            u32x4 fa[S], fb[S];
#pragma unroll
            for (int q = 0; q < S; ++q) {
                if (FULL || q < si) {
                    fa[q] = *reinterpret_cast<const u32x4*>(la + q * M::PLANE_A + oz_chunk(fr_a, kh));
                    fb[q] = *reinterpret_cast<const u32x4*>(lb + q * M::PLANE_B + oz_chunk(fr_b, kh));
                }
            }
#pragma unroll
            for (int pa = 1; pa <= S; ++pa)
#pragma unroll
                for (int qb = 1; pa + qb <= S + 1; ++qb)
                    if (FULL || pa + qb <= si + 1)
                        acc[pa + qb - 2] = __builtin_amdgcn_mfma_i32_32x32x32_i8((i32x4)fa[pa - 1], (i32x4)fb[qb - 1],
                                                                                acc[pa + qb - 2], 0, 0, 0);
        }
        // fold: orders above si + 1 are zero and leave D unchanged
#pragma unroll
        for (int t = S + 1; t >= 2; --t) {
            const double scale = ldexp(1.0, -7 * t);
#pragma unroll
            for (int v = 0; v < 16; ++v) D[v] = __builtin_fma((double)acc[t - 2][v], scale, D[v]);
        }
        if (!CHUNKS) break;
    }

    // epilogue: lane -> column (lane & 31), accumulator v -> row
    // 8 (v / 4) + 4 (lane / 32) + v % 4 of the wave's 32 x 32 block
    const int c = T.j0 + wn * 32 + (lane & 31);
    if (c >= p.n) return;
    const int32_t* ex = reinterpret_cast<const int32_t*>(smem + M::EXP + (i & 3) * M::EXP_TILE);
    const int fj = ex[BM + wn * 32 + (lane & 31)];
    double* C = wave_uniform(static_cast<double*>(p.C[T.inst]));
    const bool beta0 = p.beta0 != 0;
    const bool vec = (p.ldc % 2) == 0 && aligned16(C);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r0 = wm * 32 + 8 * q + 4 * kh;  // row within the tile
        const int row0 = T.i0 + r0;
        if (row0 >= p.m) continue;
        double val[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ei = ex[r0 + v];
            val[v] = (ei == OZ_NONFINITE || fj == OZ_NONFINITE) ? __builtin_nan("") : ldexp(D[4 * q + v], ei + fj);
        }
        double* dst = C + (long)c * p.ldc + row0;
        if (vec && row0 + 4 <= p.m) {
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                f64x2 old = {0.0, 0.0};
                if (!beta0) old = *gbl(reinterpret_cast<const f64x2*>(dst + 2 * hh));
                f64x2 res;
                res[0] = __builtin_fma(val[2 * hh], p.sc.ad, beta0 ? 0.0 : old[0] * p.sc.bd);
                res[1] = __builtin_fma(val[2 * hh + 1], p.sc.ad, beta0 ? 0.0 : old[1] * p.sc.bd);
                *gbl(reinterpret_cast<f64x2*>(dst + 2 * hh)) = res;
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

template <int S, int BM, bool CHUNKS>
__global__ void __launch_bounds__(4 * BM) ozaki_mma(const OzParams p)  // 64 * OzMma<S, BM>::NW threads
{
    using M = OzMma<S, BM>;
    __shared__ __attribute__((aligned(1024))) char smem[M::LDS_BYTES];
    const int ntiles = p.total > (long)blockIdx.x ? (int)((p.total - 1 - blockIdx.x) / gridDim.x) + 1 : 0;
    const long gend = (long)ntiles * p.nkb;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    OzStream L{0, 0, 0, nullptr, nullptr, 0, 0, 0};
    if (gend > 0) oz_stream_tile<BM>(p, L);
#pragma unroll
    for (int q = 0; q < M::AHEAD; ++q)
        if (L.g < gend) oz_issue<S, BM>(p, L, gend, smem, wave, lane);
    long g = 0;
    for (int i = 0; i < ntiles; ++i) {
        const OzTile T = oz_tile<BM>(p, i);
        const int si = oz_scalar_byte(p.sarr, T.inst);
        if (si >= S) oz_mma_tile<S, BM, true, CHUNKS>(p, T, i, S, smem, g, L, gend);
        else oz_mma_tile<S, BM, false, CHUNKS>(p, T, i, si, smem, g, L, gend);
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
// are ozaki_mma's, so the result is bit-identical to the general path.
// ---------------------------------------------------------------------------

// Slices of 16 values (v = x * 2^-e, |v| < 1) into S 16-byte chunks: slice
// p holds, for every value, its p-th base-128 digit with the value's sign
// (the digits the sequence "v *= 128; d = trunc(v); v -= d" gives, as in
// step 2 above).  Done in integers: H = floor(|v| 2^28) carries digits 1-4,
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
            // |x| 2^-e 2^28 in one scaling (where |x| 2^-e would be
            // subnormal every digit is zero either way); the remainders by
            // v_fract_f64, exact for these non-negative values below 2^28
            const double y1 = (e == OZ_NONFINITE) ? 0.0 : ldexp(fabs(x[i]), 28 - e);
            neg |= (x[i] < 0.0 ? 0xffu : 0u) << (8 * q);
            H[q] = (uint32_t)y1;
            if constexpr (S > 4) {
                const double y2 = ldexp(__builtin_amdgcn_fract(y1), 28);
                L[q] = (uint32_t)y2;
                if constexpr (S > 8) N[q] = (uint32_t)ldexp(__builtin_amdgcn_fract(y2), 7);
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

// Reductions over the 32 row (column) owners (lanes l and l ^ 32 hold the
// same values) without LDS-routed shuffles: within each 16-lane DPP row by
// quad swaps (xor 1, xor 2) and row rotations by 4 and 8 (every lane then
// holds its row's total), across the two rows of a half by ds_swizzle
// (xor 16).  The joins of the two k halves (xor 32) use v_permlane32_swap.
template <class Op>
__device__ __forceinline__ uint32_t oz_row16(uint32_t v, Op op)
{
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));  // row_ror:4
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));  // row_ror:8
    return op(v, (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F));                    // xor 16
}
__device__ __forceinline__ float oz_sum32(float v)
{
    return __builtin_bit_cast(float, oz_row16(__builtin_bit_cast(uint32_t, v), [](uint32_t a, uint32_t b) {
        return __builtin_bit_cast(uint32_t, __builtin_bit_cast(float, a) + __builtin_bit_cast(float, b));
    }));
}
__device__ __forceinline__ int oz_max32(int v)
{
    return (int)oz_row16((uint32_t)v, [](uint32_t a, uint32_t b) { return (uint32_t)max((int)a, (int)b); });
}
// The value of lane l ^ 32.
__device__ __forceinline__ uint32_t oz_other_half(uint32_t v)
{
    const auto pr = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    // pr[0]: the lower half's values in both halves, pr[1]: the upper half's
    return (threadIdx.x & 32) ? pr[0] : pr[1];
}
__device__ __forceinline__ float oz_other_half(float v)
{
    return __builtin_bit_cast(float, oz_other_half(__builtin_bit_cast(uint32_t, v)));
}

// Bound at S slices for the fused one-wave form: lane l owns row l % 32 of
// op(A) and column l % 32 of op(B) over the k half l / 32 (the statistics of
// both halves are joined first).  The fused kernel's time is set by its
// loads and its slicing, not by the slice count, so its automatic choice is
// always the most slices and only the bound at S is evaluated here; the
// general path chooses per instance (ozaki_choose).
template <int S>
__device__ __forceinline__ float oz_bound(int ea, int eb, OzStats sa, OzStats sb, const uint32_t (&dsum)[S])
{
    if (__any(ea == OZ_NONFINITE || eb == OZ_NONFINITE)) return __builtin_huge_valf();
    // join the two k halves
    float a1 = sa.l1(), a2 = sa.l2(), b1 = sb.l1(), b2 = sb.l2();
    a1 += oz_other_half(a1);
    a2 += oz_other_half(a2);
    b1 += oz_other_half(b1);
    b2 += oz_other_half(b2);
    uint32_t Sa = 0;  // digit sum S_i(S)
#pragma unroll
    for (int p = 0; p < S; ++p) Sa += dsum[p];
    Sa += oz_other_half(Sa);
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


// Exponent of the max |.| over both k halves (lanes l and l ^ 32), as
// exp_of_max gives it: from the largest biased exponent field E of the
// values (the high words of |x| compare as the values do), e = E - 1022 for
// a normal maximum, OZ_NONFINITE for E = 0x7ff (inf / nan); a maximum that
// is zero or subnormal (E = 0) goes through frexp.
__device__ __forceinline__ int oz_exp16(const double* x)
{
    uint32_t hmax = 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) hmax = max(hmax, (uint32_t)(__builtin_bit_cast(uint64_t, x[i]) >> 32) & 0x7fffffffu);
    hmax = max(hmax, oz_other_half(hmax));
    const int E = (int)(hmax >> 20);
    if (E == 0x7ff) return OZ_NONFINITE;
    if (E > 0) return E - 1022;
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) mx = fmax(mx, fabs(x[i]));
    mx = fmax(mx, __shfl_xor(mx, 32));
    return exp_of_max(mx, false);
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

// One instance of the fused form once its operands are in registers (lane l:
// row l % 32 of op(A), column l % 32 of op(B), k half l / 32): exponents,
// slices, the triangle of slice products on v_mfma_i32_32x32x32_i8, the fold
// (smallest order first), the epilogue into C and the error report.
// before_epilogue() runs between the products and the first access to C.
// xch: the wave's 32 ints of LDS for the row exponents.
template <int S, class Hook>
__device__ __forceinline__ void oz_small_instance(const OzParams& p, long inst, const double (&xa)[16],
                                                  const double (&xb)[16], double* C, int lane, int* xch,
                                                  Hook&& before_epilogue)
{
    const int own = lane & 31, kh = lane >> 5;
    const bool beta0 = p.beta0 != 0;
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
    before_epilogue();

    // epilogue: C = fma(D * 2^(e_i + f_j), alpha, fl(beta * C)).  Lane
    // (own, kh) holds rows 8 q + 4 kh .. + 3 of column own, q = 0..3.  With
    // all 32 rows inside the instance and C 16-byte aligned, they go out as
    // 16-byte pairs, beta == 0 and beta != 0 decided once for the wave and
    // the old values read before the first store: a load issued
    // conditionally between the stores puts a vmcnt wait at its join, and
    // on gfx9 that wait also waits for every store issued before it.
    const int c = own;
    // row exponents through LDS (lane `row` holds that row's; lanes l and
    // l + 32 write the same value): four consecutive rows per 16-byte read
    xch[own] = ea;
    int eir[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const i32x4 r = *reinterpret_cast<const i32x4*>(xch + 8 * q + 4 * kh);
#pragma unroll
        for (int j = 0; j < 4; ++j) eir[4 * q + j] = r[j];
    }
    if (c < p.n) {
        double val[16];
#pragma unroll
        for (int v = 0; v < 16; ++v)
            val[v] = (eir[v] == OZ_NONFINITE || eb == OZ_NONFINITE) ? __builtin_nan("") : ldexp(D[v], eir[v] + eb);
        double* col = C + (long)c * p.ldc + 4 * kh;
        if ((p.ldc % 2) == 0 && aligned16(C) && p.m >= 32) {
            if (beta0) {
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    f64x2 res;
                    res[0] = __builtin_fma(val[2 * h], p.sc.ad, 0.0);
                    res[1] = __builtin_fma(val[2 * h + 1], p.sc.ad, 0.0);
                    *gbl(reinterpret_cast<f64x2*>(col + 8 * (h >> 1) + 2 * (h & 1))) = res;
                }
            } else {
                f64x2 old[8];
#pragma unroll
                for (int h = 0; h < 8; ++h) old[h] = *gbl(reinterpret_cast<const f64x2*>(col + 8 * (h >> 1) + 2 * (h & 1)));
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    f64x2 res;
                    res[0] = __builtin_fma(val[2 * h], p.sc.ad, old[h][0] * p.sc.bd);
                    res[1] = __builtin_fma(val[2 * h + 1], p.sc.ad, old[h][1] * p.sc.bd);
                    *gbl(reinterpret_cast<f64x2*>(col + 8 * (h >> 1) + 2 * (h & 1))) = res;
                }
            }
        } else {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = 8 * (v >> 2) + 4 * kh + (v & 3);
                if (row < p.m) {
                    double* dst = col + 8 * (v >> 2) + (v & 3);
                    const double t0 = beta0 ? 0.0 : (*gbl(dst)) * p.sc.bd;
                    *gbl(dst) = __builtin_fma(val[v], p.sc.ad, t0);
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

// Persistent: wave w of the grid computes instances w, w + W, w + 2W, ...
// (W = waves in the grid), loading each instance's operands straight into
// registers (instances smaller than 32 x 32 x 32: guarded loads).
template <int S>
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
    // op(A) row own is K-contiguous when A is transposed; op(B) column own
    // when B is not
    const OzLane LA = oz_lane(p.lda, p.ta != 0, p.m, p.k, own, kh);
    const OzLane LB = oz_lane(p.ldb, p.tb == 0, p.n, p.k, own, kh);
    __shared__ __attribute__((aligned(16))) int xch[4 * 32];
    for (; inst < p.batch; inst += W) {
        double xa[16], xb[16];
        oz_small_load<false>(p.A[inst], LA, xa);
        oz_small_load<false>(p.B[inst], LB, xb);
        oz_small_instance<S>(p, inst, xa, xb, wave_uniform(static_cast<double*>(p.C[inst])), lane, xch + 32 * wave,
                             [] {});
    }
}

// ---------------------------------------------------------------------------
// Whole 32 x 32 x 32 instances: each wave owns a 16 KiB LDS image of one
// instance (op(A), then op(B), 8 KiB each) and fills it with
// global_load_lds_dwordx4 (1 KiB per wave instruction, no registers): as soon
// as it has read the current instance out of the image into registers, it
// requests the next instance's operands into it, so their loads are in
// flight while the current instance is sliced and multiplied (the register
// form above waits for every instance's loads in turn, at the two waves per
// SIMD its registers allow).
//
// Image of one operand, 16-byte chunks (two doubles):
//   M/N-contiguous (element (r, k) at r + k ld): [k][row], chunk c of k-row
//     kk at slot 16 kk + c; the lane reads its 16 values with ds_read_b64;
//   K-contiguous (element (r, k) at r ld + k): [row][k], chunk c of row r at
//     slot 16 r + (c ^ (r & 15)), so the 32 rows a read instruction covers
//     fall on 16 different 16-byte bank groups; ds_read_b128.
// 16-byte loads need 16-byte aligned chunks (even leading dimension, A or B
// 16-byte aligned); otherwise the operand's image is filled with 4-byte
// loads (global_load_lds_dword, 256 bytes per wave instruction) into the
// same layout.
//
// Pointer words are read two instances ahead (lane-packed in one VGPR) and
// taken into use once they have landed, before the epilogue.
// Waits: the image's loads are not tracked by the compiler.  Behind the
// loads of an instance's image the wave issues at least nine more vector
// memory instructions before it reads the image (the pointer words of the
// instance after, and at least eight C stores of a 32 x 32 fp64 block), so
// vmcnt(9) there means the image is complete; the first image is waited
// for in full.
// ---------------------------------------------------------------------------
constexpr int OZ_IMG = 32 * 32 * 8;  // one operand

// The lane index goes through an empty asm in these helpers: left visible,
// the compiler hoists every lane-dependent address out of the instance loop
// and keeps them all live in registers across it.
__device__ __forceinline__ void oz_dma_operand(const double* X, long ld, bool kc, char* img, int lane_in)
{
    int lane = lane_in;
    __asm__ volatile("" : "+v"(lane));
    const char* base = reinterpret_cast<const char*>(X);
    if (((ld % 2) == 0) && aligned16(X)) {
        // slot 64 q + lane: row (KC) or k (M/N) 4 q + (lane >> 4), chunk
        // lane & 15 (KC: stored at chunk c ^ (row & 15), and row & 15 =
        // 4 (q & 3) + (lane >> 4)); the lane's address advances by 4 ld
        // elements per instruction
        const int r0 = lane >> 4, c = lane & 15;
        const char* lane_base = base + 8 * ((long)r0 * ld);
        const long step = 32 * ld;  // bytes per 4 rows
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int cg = kc ? (c ^ (4 * (q & 3) + r0)) : c;
            __builtin_amdgcn_global_load_lds(lane_base + q * step + 16 * cg, (lds_void_t*)(img + 1024 * q), 16, 0, 0);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int word = 64 * q + lane;  // 4-byte word of the image
            const int slot = word >> 2, hi = slot >> 4, c = slot & 15;
            const long off = (long)hi * ld + 2 * (kc ? (c ^ (hi & 15)) : c);
            __builtin_amdgcn_global_load_lds(base + 8 * off + 4 * (word & 3), (lds_void_t*)(img + 256 * q), 4, 0, 0);
        }
    }
}

__device__ __forceinline__ void oz_img_read(const char* img, bool kc, int lane_in, double (&x)[16])
{
    int lane = lane_in;
    __asm__ volatile("" : "+v"(lane));
    const int own = lane & 31, kh = lane >> 5;
    if (kc) {
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const int c = 8 * kh + h;
            const f64x2 v = *reinterpret_cast<const f64x2*>(img + 256 * own + 16 * (c ^ (own & 15)));
            x[2 * h] = v[0];
            x[2 * h + 1] = v[1];
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = *reinterpret_cast<const double*>(img + 256 * (16 * kh + q) + 8 * own);
    }
}

// Pointer words of instance `inst`, lane-packed: lanes 0-1 A, 2-3 B, 4-5 C
// (low word first); the other lanes read C's high word again.
__device__ __forceinline__ uint32_t oz_ptr_word(const OzParams& p, long inst, int lane)
{
    const int w = lane < 5 ? lane : 5;
    const void* const* arr = w < 2 ? p.A : w < 4 ? p.B : const_cast<const void* const*>(p.C);
    return *gbl(reinterpret_cast<const uint32_t*>(arr + inst) + (w & 1));
}
__device__ __forceinline__ const void* oz_ptr_of(uint32_t v, int which)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)v, 2 * which);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)v, 2 * which + 1);
    return reinterpret_cast<const void*>(((uint64_t)hi << 32) | lo);
}
// A copy of v once its load has landed (carried across iterations with no
// load in flight behind it)
__device__ __forceinline__ uint32_t oz_settle(uint32_t v)
{
    uint32_t r;
    __asm__ volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

template <int S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) ozaki_small_full(const OzParams p)
{
    __shared__ __attribute__((aligned(1024))) char smem[4 * 2 * OZ_IMG];
    __shared__ __attribute__((aligned(16))) int xch[4 * 32];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long W = (long)gridDim.x * 4;
    long inst = (long)blockIdx.x * 4 + wave;
    if (inst >= p.batch) return;
    const bool akc = p.ta != 0, bkc = p.tb == 0;
    char* img = smem + wave * 2 * OZ_IMG;
    auto dma = [&](uint32_t pw) {
        oz_dma_operand(static_cast<const double*>(oz_ptr_of(pw, 0)), p.lda, akc, img, lane);
        oz_dma_operand(static_cast<const double*>(oz_ptr_of(pw, 1)), p.ldb, bkc, img + OZ_IMG, lane);
    };
    uint32_t pw = oz_settle(oz_ptr_word(p, inst, lane));
    uint32_t pwn = oz_settle(oz_ptr_word(p, inst + W < p.batch ? inst + W : inst, lane));
    dma(pw);
    __asm__ __volatile__("s_waitcnt vmcnt(0)" ::: "memory");
    for (; inst < p.batch; inst += W) {
        const long nxt = inst + W, after = inst + 2 * W;
        double xa[16], xb[16];
        __asm__ __volatile__("s_waitcnt vmcnt(9)" ::: "memory");
        oz_img_read(img, akc, lane, xa);
        oz_img_read(img + OZ_IMG, bkc, lane, xb);
        // read out before the next instance's loads overwrite the image
        __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory");
        double* C = static_cast<double*>(const_cast<void*>(oz_ptr_of(pw, 2)));
        if (nxt < p.batch) dma(pwn);
        const uint32_t t = oz_ptr_word(p, after < p.batch ? after : inst, lane);
        uint32_t pwa = 0u;
        oz_small_instance<S>(p, inst, xa, xb, C, lane, xch + 32 * wave, [&] { pwa = oz_settle(t); });
        pw = pwn;
        pwn = pwa;
    }
}

template <int S, bool FULL>
static hipError_t run_ozaki_small_k(const OzParams& p, hipStream_t st)
{
    auto kern = FULL ? ozaki_small_full<S> : ozaki_small<S>;
    const int occ = kernel_occupancy(kern, 256);
    const long blocks = ((long)p.batch + 3) / 4;
    long grid = (long)occ * device_cu_count();
    if (const char* e = getenv("SSBLAS_MAX_GRID")) {
        const long cap = atol(e);
        if (cap >= 1 && cap < grid) grid = cap;
    }
    if (blocks < grid) grid = blocks;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, p);
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
template <int S, int BM, bool CHUNKS>
static hipError_t run_ozaki_mma_k(const OzParams& p0, hipStream_t st)
{
    using M = OzMma<S, BM>;
    auto kern = ozaki_mma<S, BM, CHUNKS>;
    const int occ = kernel_occupancy(kern, 64 * M::NW);
    const long slots = (long)occ * device_cu_count();
    OzParams p = p0;
    p.tiles_m = (p.m + BM - 1) / BM;
    p.tiles_n = (p.n + OZ_ROWS - 1) / OZ_ROWS;
    const long tpi = (long)p.tiles_m * p.tiles_n;
    // the kernel's tile arithmetic is 32-bit: launches of fewer than 2^31 tiles
    const long per_launch = ((1L << 31) - 1) / tpi;
    for (long b0 = 0; b0 < p0.batch; b0 += per_launch) {
        p.batch = (int)std::min<long>(per_launch, p0.batch - b0);
        p.A = p0.A + b0;
        p.B = p0.B + b0;
        p.C = p0.C + b0;
        p.wa = p0.wa + b0 * p0.inst_a;
        p.wb = p0.wb + b0 * p0.inst_b;
        p.ea = p0.ea + b0 * p0.m;
        p.eb = p0.eb + b0 * p0.n;
        p.sarr = p0.sarr + b0;
        p.total = (long)p.batch * tpi;
        // a multiple of 8 workgroups keeps workgroup b on XCD b % 8 for every
        // tile it takes (xcd_remap)
        const long grid = p.total <= slots ? p.total : (slots / 8) * 8;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * M::NW), 0, st, p);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int S>
static hipError_t run_ozaki_mma_s(const OzParams& p, hipStream_t st)
{
    const bool chunks = p.nkb > OZ_KCHUNK / OZ_KB;
    if (ozaki_tile_rows(p.m) == 128)
        return chunks ? run_ozaki_mma_k<S, 128, true>(p, st) : run_ozaki_mma_k<S, 128, false>(p, st);
    return chunks ? run_ozaki_mma_k<S, 64, true>(p, st) : run_ozaki_mma_k<S, 64, false>(p, st);
}

static hipError_t run_ozaki_mma(const OzParams& p, hipStream_t st)
{
    switch (p.planes) {
    case 3: return run_ozaki_mma_s<3>(p, st);
    case 4: return run_ozaki_mma_s<4>(p, st);
    case 5: return run_ozaki_mma_s<5>(p, st);
    case 6: return run_ozaki_mma_s<6>(p, st);
    case 7: return run_ozaki_mma_s<7>(p, st);
    case 8: return run_ozaki_mma_s<8>(p, st);
    default: return run_ozaki_mma_s<9>(p, st);
    }
}

// Exponents, slices and statistics of one operand's rows (op(A)) or columns
// (op(B)).  kc: element (row, k) at row * ld + k.
static hipError_t prep_operand(const void* const* X, long ld, int rows, int k, bool kc, int batch, int planes,
                               long rows_pad, long inst_bytes, int8_t* w, int32_t* e, OzRow* stats, hipStream_t st)
{
    OzPrep q;
    q.X = X;
    q.ld = ld;
    q.rows = rows;
    q.k = k;
    q.nrb = (int)(rows_pad / OZ_ROWS);
    q.planes = planes;
    q.rows_pad = rows_pad;
    q.inst_bytes = inst_bytes;
    q.w = w;
    q.e = e;
    q.st = stats;
    const dim3 grid((unsigned)((long)batch * q.nrb));
    // the operand stays in registers between the passes up to K = 256
    auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, q); };
    if (k <= 64) kc ? launch(ozaki_prep<true, 1>) : launch(ozaki_prep<false, 1>);
    else if (k <= 128) kc ? launch(ozaki_prep<true, 2>) : launch(ozaki_prep<false, 2>);
    else if (k <= 256) kc ? launch(ozaki_prep<true, 4>) : launch(ozaki_prep<false, 4>);
    else kc ? launch(ozaki_prep<true, 0>) : launch(ozaki_prep<false, 0>);
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

    p.wa = reinterpret_cast<int8_t*>(ws + L.off_a);
    p.wb = reinterpret_cast<int8_t*>(ws + L.off_b);
    p.ea = reinterpret_cast<int32_t*>(ws + L.off_e);
    p.eb = p.ea + (size_t)g.batch * m;
    p.ra = reinterpret_cast<OzRow*>(ws + L.off_r);
    p.rb = p.ra + (size_t)g.batch * m;
    p.rows_pad_a = L.rows_pad_a;
    p.rows_pad_b = L.rows_pad_b;
    p.inst_a = (long)L.inst_a;
    p.inst_b = (long)L.inst_b;
    p.nkb = (int)L.nkb;
    p.planes = s;

    // op(A) row i is K-contiguous when A is transposed; op(B) column j when
    // B is not
    err = prep_operand(g.A, g.lda, m, k, g.ta != 0, g.batch, s, L.rows_pad_a, (long)L.inst_a, p.wa, p.ea, p.ra, st);
    if (err != hipSuccess) return err;
    err = prep_operand(g.B, g.ldb, n, k, g.tb == 0, g.batch, s, L.rows_pad_b, (long)L.inst_b, p.wb, p.eb, p.rb, st);
    if (err != hipSuccess) return err;
    // error bound and slice count per instance
    hipLaunchKernelGGL(ozaki_choose, dim3((unsigned)g.batch), dim3(256), 0, st, p);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    return run_ozaki_mma(p, st);
}

}  // namespace ssb
