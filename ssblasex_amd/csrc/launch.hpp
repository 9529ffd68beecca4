// launch.hpp -- internal interface between the host API layer (api.cpp) and
// the kernel translation units.  Plain C++ types only, so api.cpp compiles
// with the host compiler.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace ssb {

enum ArithKind : int {
    K_I8 = 0,
    K_F32 = 1,
    K_F64 = 2,
    K_BF16X3 = 3,
    K_BF16X6 = 4,
};

// Tile configurations.  DIRECT_*: one wave per block, no LDS.  TILE_*: LDS
// double-buffered workgroup tiles (k-depth per form in launch_impl.hpp).
enum TileConfig : int {
    CFG_DIRECT_16 = 0,     // wave per 16x16
    CFG_DIRECT_32 = 1,     // wave per 32x32
    CFG_TILE_64 = 2,       // 64x64, 4 waves of 32x32
    CFG_TILE_128 = 3,      // 128x128, 4 waves of 64x64
    CFG_TILE_128x64 = 4,   // 128x64 (M x N), 4 waves of 64x32
    CFG_TILE_64x128 = 5,   // 64x128, 4 waves of 32x64
    CFG_TILE_128_K2 = 6,   // 128x128, 4 waves, twice the k-depth of TILE_128
    CFG_TILE_128_W8 = 7,   // 128x128, 8 waves of 32x64
    CFG_TILE_256x128 = 8,  // 256x128, 8 waves of 64x64
    // the same shapes with loads issued two k-blocks ahead
    CFG_TILE_64_S2 = 9,
    CFG_TILE_128_S2 = 10,
    CFG_TILE_128_K2_S2 = 11,
    CFG_TILE_128_W8_S2 = 12,
    CFG_TILE_256x128_S2 = 13,
    // int8 only (other forms use TILE_128): direct-to-LDS loads, S stages
    CFG_DMA_64 = 14,        // 64x64, k 64, 4 stages
    CFG_DMA_128 = 15,       // 128x128, k 64, 4 stages
    CFG_DMA_128_K2 = 16,    // 128x128, k 128, 3 stages
    CFG_DMA_256x128 = 17,   // 256x128, 8 waves, k 64, 3 stages
    CFG_DMA_128_S3 = 18,    // 128x128, k 64, 3 stages
    // fp32 / fp64 / bf16 forms with K-contiguous operands (A^T, B): direct-to-
    // LDS loads of the input elements, bf16 split in registers (int8 uses
    // DMA_128; other layouts the TILE config of the same shape)
    CFG_KC_256x128 = 19,    // 256x128, 8 waves, 128 bytes of k per row, 3 stages
    CFG_KC_128 = 20,        // 128x128, 4 waves, 128 bytes of k per row, 2 stages
    // int8 only: 256x256 direct-to-LDS tiles (one per CU for a batch of 1024
    // 128x128 tiles), k 64, 4 stages
    CFG_DMA_256 = 21,       // 8 waves of 128x64
    CFG_DMA_256_W16 = 22,   // 16 waves of 64x64
    CFG_DMA_256_K2 = 23,    // 8 waves of 128x64, k 128, 2 stages
    // int8 and fp32 (other forms: TILE_128): 128x128, 4 waves of 64x64, k-blocks
    // direct to LDS through range-checked buffer loads (any layout, leading
    // dimension and alignment; the grouped kernel's tile), three stages
    CFG_GDMA_128 = 24,
    CFG_GDMA_64 = 25,  // the same with 64x64 tiles (four waves of 32x32; other forms: TILE_64)
    CFG_GDMA_256x128 = 26,  // 256x128, 8 waves of 64x64 (other forms: TILE_256x128)
    CFG_COUNT = 27
};

const char* config_name(int cfg);
void config_tile(int cfg, int* tm, int* tn);

struct GemmProblem {
    int arith;             // ArithKind
    bool ta, tb;           // op(A) == A^T, op(B) == B^T
    int m, n, k;
    long lda, ldb, ldc;
    int batch;
    const void* const* A;  // device-accessible arrays of device pointers
    const void* const* B;
    void* const* C;
    float af, bf;          // scalars in every representation
    double ad, bd;
    int32_t ai, bi;
    bool beta0;
};

// Enqueue one batched GEMM with the given tile config on `stream`.
hipError_t launch_gemm(const GemmProblem& g, int cfg, hipStream_t stream);

// Default tile config for a problem (size-bucket heuristic, overridable by
// the autotune table).
int default_config(const GemmProblem& g);

// Per-arith entry points (one translation unit each).
hipError_t launch_i8(const GemmProblem& g, int cfg, hipStream_t stream);
hipError_t launch_f32(const GemmProblem& g, int cfg, hipStream_t stream);
hipError_t launch_f64(const GemmProblem& g, int cfg, hipStream_t stream);
hipError_t launch_bf16x3(const GemmProblem& g, int cfg, hipStream_t stream);
hipError_t launch_bf16x6(const GemmProblem& g, int cfg, hipStream_t stream);

// Number of compute units of the current device (cached per device).
int device_cu_count();

}  // namespace ssb
