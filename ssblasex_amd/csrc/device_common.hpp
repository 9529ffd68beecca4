// device_common.hpp -- shared device-side vocabulary for the gfx950 kernels:
// vector types, the per-type arithmetic policies (MFMA instruction, LDS
// planes, fused alpha/beta epilogue) and the byte/word transposes used when
// an operand's contiguous dimension is M or N rather than K.
//
// MFMA shape used everywhere: 16x16 output tiles.  Output layout of
// i32_16x16x64_i8, f32_16x16x4f32 and f32_16x16x32_bf16: lane l,
// accumulator slot v holds row 4*(l/16)+v, column l%16 of the 16x16 tile.
// f64_16x16x4f64 instead places MFMA row g+4v in lane group g, slot v; its
// A fragments are therefore loaded with rows permuted (frag_row below) so
// that every form ends with the same lane -> (4 consecutive rows, column)
// layout and shares one epilogue.
//
// Operand layout convention: for one "k-step" a lane of the wave supplies
// row (l%16) of the A tile and column (l%16) of the B tile, and lane group
// g = l/16 supplies the 16-byte chunk of K elements starting at
// g * (16 / element bytes).  A and B use the identical lane-group -> k
// assignment, so the hardware's internal ordering of the k elements inside
// an operand register never matters: every product a(i,k)*b(k,j) meets.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssb {

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Identifiers of the arithmetic forms (shared with the host launcher).
enum ArithId : int {
    ARITH_I8 = 0,       // int8 x int8 -> int32, exact modulo 2^32
    ARITH_F32 = 1,      // native fp32 MFMA
    ARITH_F64 = 2,      // native fp64 MFMA
    ARITH_BF16X3 = 3,   // fp32 via bf16 hi/lo, 3 products
    ARITH_BF16X6 = 4,   // fp32 via bf16 hi/mid/lo, 6 products
};

// Scalars of one call, every representation filled by the host.
// Global-address-space views of generic pointers.  Operand and result
// pointers come out of pointer arrays, so the compiler cannot tell they are
// global and would use FLAT instructions; a FLAT access counts against the
// LDS wait counter too, so every wait for an LDS read would also wait for
// the global loads and stores in flight.
#define SSB_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const SSB_GLOBAL T* gbl(const T* p)
{
    return (const SSB_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ SSB_GLOBAL T* gbl(T* p)
{
    return (SSB_GLOBAL T*)p;
}

struct Scalars {
    float af, bf;
    double ad, bd;
    int32_t ai, bi;
};

// ---------------------------------------------------------------------------
// bf16 split of two fp32 values (round to nearest even at every cut).
// hi = RN_bf16(x), r1 = x - hi (exact in fp32), mid = RN_bf16(r1), ...
// Each cut is one v_cvt_pk_bf16_f32 (two values, hardware RNE); plane p of
// the pair comes back as one dword, x0 in the low half.  A non-finite value
// keeps its hi part (inf stays inf, NaN stays NaN); its tails are NaN, so
// the C entries it reaches are NaN.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t bf16_pack_rn(float x0, float x1)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){x0, x1}, bf16x2));
}

template <int NP>
__device__ __forceinline__ void bf16_split2(float x0, float x1, uint32_t (&parts)[NP])
{
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const uint32_t h = bf16_pack_rn(x0, x1);
        parts[p] = h;
        if (p + 1 < NP) {
            x0 = x0 - __uint_as_float(h << 16);  // exact (Sterbenz-type cancellation)
            x1 = x1 - __uint_as_float(h & 0xffff0000u);
        }
    }
}

// ---------------------------------------------------------------------------
// Byte transpose of a 4x4 block held in four dwords:
//   in[r] = (x[r][0], x[r][1], x[r][2], x[r][3])  (byte 0 = lowest)
//   out[c] = (x[0][c], x[1][c], x[2][c], x[3][c])
// v_perm_b32(hi, lo, sel): selector byte s picks byte s of {hi:lo}
// (0..3 from lo, 4..7 from hi).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void transpose4x4_bytes(const uint32_t in[4], uint32_t out[4])
{
    // t0 = (in0.b0, in1.b0, in0.b1, in1.b1), t1 = (in0.b2, in1.b2, in0.b3, in1.b3)
    uint32_t t0 = __builtin_amdgcn_perm(in[1], in[0], 0x05010400u);
    uint32_t t1 = __builtin_amdgcn_perm(in[1], in[0], 0x07030602u);
    uint32_t t2 = __builtin_amdgcn_perm(in[3], in[2], 0x05010400u);
    uint32_t t3 = __builtin_amdgcn_perm(in[3], in[2], 0x07030602u);
    // out0 = (t0.b0, t0.b1, t2.b0, t2.b1), out1 = (t0.b2, t0.b3, t2.b2, t2.b3)
    out[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
    out[1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
    out[2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
    out[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

// ---------------------------------------------------------------------------
// Arithmetic policies.
//   In      global element type of A and B
//   Out     element type of C
//   E       bytes per In element
//   NPLANE  LDS planes per operand (1; 2 or 3 for the bf16 splits)
//   LDS_E   bytes per LDS element
//   KSTEP   K elements consumed by one MFMA "k-step" (4 lane groups x 16 B
//           of LDS elements)
//   Acc     16x16 accumulator fragment
// ---------------------------------------------------------------------------
struct ArithI8 {
    using In = int8_t;
    using Out = int32_t;
    using Acc = i32x4;
    static constexpr int E = 1, NPLANE = 1, LDS_E = 1, KSTEP = 64;
    static constexpr int ID = ARITH_I8;
    static constexpr bool ROW_PERM = false;
    static constexpr bool SWZ = false;
    __device__ static __forceinline__ void mma(Acc& c, const u32x4* a, const u32x4* b)
    {
        c = __builtin_amdgcn_mfma_i32_16x16x64_i8((i32x4)a[0], (i32x4)b[0], c, 0, 0, 0);
    }
    // 16 input bytes -> plane chunks (identity)
    __device__ static __forceinline__ void to_planes(const u32x4& in, u32x4* planes) { planes[0] = in; }
};

struct ArithF32 {
    using In = float;
    using Out = float;
    using Acc = f32x4;
    static constexpr int E = 4, NPLANE = 1, LDS_E = 4, KSTEP = 16;
    static constexpr int ID = ARITH_F32;
    static constexpr bool ROW_PERM = false;
    static constexpr bool SWZ = false;
    __device__ static __forceinline__ void mma(Acc& c, const u32x4* a, const u32x4* b)
    {
        f32x4 fa = (f32x4)a[0], fb = (f32x4)b[0];
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[0], fb[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[1], fb[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[2], fb[2], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[3], fb[3], c, 0, 0, 0);
    }
    __device__ static __forceinline__ void to_planes(const u32x4& in, u32x4* planes) { planes[0] = in; }
};

struct ArithF64 {
    using In = double;
    using Out = double;
    using Acc = f64x4;
    static constexpr int E = 8, NPLANE = 1, LDS_E = 8, KSTEP = 8;
    static constexpr int ID = ARITH_F64;
    static constexpr bool ROW_PERM = true;
    static constexpr bool SWZ = false;
    __device__ static __forceinline__ void mma(Acc& c, const u32x4* a, const u32x4* b)
    {
        f64x2 da = __builtin_bit_cast(f64x2, a[0]);
        f64x2 db = __builtin_bit_cast(f64x2, b[0]);
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(da[0], db[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(da[1], db[1], c, 0, 0, 0);
    }
    __device__ static __forceinline__ void to_planes(const u32x4& in, u32x4* planes) { planes[0] = in; }
};

// fp32 operands held as NP bf16 planes; products of combined order <= ORD
// (order of hi = 0, mid = 1, lo = 2).  X3: NP = 2, ORD = 1; X6: NP = 3, ORD = 2.
template <int NP, int ORD, int IDV>
struct ArithBF16Split {
    using In = float;
    using Out = float;
    using Acc = f32x4;
    static constexpr int E = 4, NPLANE = NP, LDS_E = 2, KSTEP = 32;
    static constexpr int ID = IDV;
    static constexpr bool ROW_PERM = false;
    // LDS rows unpadded, 16-byte chunks XOR-swizzled (lds_row_off): three
    // planes of a 256x128 tile then fit the 160 KiB
    static constexpr bool SWZ = true;
    __device__ static __forceinline__ void mma(Acc& c, const u32x4* a, const u32x4* b)
    {
        // Smallest terms first so they are not absorbed by the large ones.
#pragma unroll
        for (int o = ORD; o >= 0; --o) {
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                int q = o - p;
                if (q < 0 || q >= NP) continue;
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[p]),
                                                            __builtin_bit_cast(bf16x8, b[q]), c, 0,
                                                            0, 0);
            }
        }
    }
    // 16 input bytes = 4 floats -> NP plane half-chunks of 4 bf16 (8 bytes)
    // returned in the low two dwords of planes[p].
    __device__ static __forceinline__ void to_planes(const u32x4& in, u32x4* planes)
    {
        uint32_t lo[NP], hi[NP];
        bf16_split2<NP>(__uint_as_float(in.x), __uint_as_float(in.y), lo);
        bf16_split2<NP>(__uint_as_float(in.z), __uint_as_float(in.w), hi);
#pragma unroll
        for (int p = 0; p < NP; ++p) planes[p] = u32x4{lo[p], hi[p], 0u, 0u};
    }
};
using ArithBF16X3 = ArithBF16Split<2, 1, ARITH_BF16X3>;
using ArithBF16X6 = ArithBF16Split<3, 2, ARITH_BF16X6>;

// ---------------------------------------------------------------------------
// Fused alpha/beta epilogue, following the reference's order of rounding:
//   fp:  t = (beta == 0) ? 0 : fl(beta * c_old);  c = fma(acc, alpha, t)
//        (CScale/Xgemm_scale_kernel.hpp:147-189 then the FMAD epilogue of
//         GemmCore/Xgemm_kernel_sve_v4x5_1.hpp:93-127)
//   int: c = alpha*acc + [beta != 0] * beta * c_old   modulo 2^32
// ---------------------------------------------------------------------------
__device__ __forceinline__ float epi(float acc, float c_old, const Scalars& s, bool beta0)
{
    float t = beta0 ? 0.0f : c_old * s.bf;
    return __builtin_fmaf(acc, s.af, t);
}
__device__ __forceinline__ double epi(double acc, double c_old, const Scalars& s, bool beta0)
{
    double t = beta0 ? 0.0 : c_old * s.bd;
    return __builtin_fma(acc, s.ad, t);
}
__device__ __forceinline__ int32_t epi(int32_t acc, int32_t c_old, const Scalars& s, bool beta0)
{
    uint32_t r = (uint32_t)s.ai * (uint32_t)acc;
    if (!beta0) r += (uint32_t)s.bi * (uint32_t)c_old;
    return (int32_t)r;
}

// Byte offset of (row, kbyte) in an LDS plane of pitch rowb.  Padded rows
// (SWZ false) carry 16 spare bytes; swizzled rows (SWZ true) are unpadded and
// 16-byte chunk c of row r sits at chunk c ^ ((r >> 1) & (chunks - 1)), so
// the 8 rows one ds_read_b128 cycle serves land on distinct banks.
template <class AR>
__device__ __forceinline__ int lds_row_off(int row, int kbyte, int rowb)
{
    if constexpr (AR::SWZ) {
        const int mask = (rowb >> 4) - 1;
        return row * rowb + ((((kbyte >> 4) ^ ((row >> 1) & mask)) << 4) | (kbyte & 15));
    } else {
        return row * rowb + kbyte;
    }
}
template <class AR>
constexpr int lds_row_pitch(int bk)
{
    return bk * AR::LDS_E + (AR::SWZ ? 0 : 16);
}

// Row of the 16-row A fragment that MFMA input lane r (= lane % 16) must
// carry so that output lane group g, slot v holds tile row 4g + v.
template <class AR>
__device__ __forceinline__ int frag_row(int r)
{
    if constexpr (AR::ROW_PERM) return 4 * (r & 3) + (r >> 2);
    else return r;
}

template <class T>
__device__ __forceinline__ bool aligned16(const T* p)
{
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

// Buffer descriptor word 3 for gfx9-family targets: 32-bit data format,
// raw (untyped) addressing.
constexpr int GFX9_BUFFER_RSRC_WORD3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, GFX9_BUFFER_RSRC_WORD3);
}

// s_waitcnt vmcnt(N) leaving expcnt / lgkmcnt at their maxima (gfx9 encoding:
// vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt_hi[15:14]).
template <int N>
__device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (7 << 4) | (0xF << 8));
}

// Map a linear slot t in [0, total) to a work id so that the work ids a
// hardware XCD sees (slots t with equal t % 8, because workgroups are
// dispatched round-robin over the 8 XCDs) form one contiguous range:
// neighbouring tiles of one instance then share that XCD's L2.  Bijective
// for every total.
__device__ __forceinline__ long xcd_remap(long t, long total)
{
    long q = total >> 3, r = total & 7;
    long x = t & 7, local = t >> 3;
    long start = x * q + (x < r ? x : r);
    return start + local;
}

}  // namespace ssb
