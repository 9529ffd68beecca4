// dispatch.cpp -- tile-config choice and the per-form launch switch.
#include <stdlib.h>
#include <string.h>

#include <mutex>

#include "api_internal.hpp"
#include "launch.hpp"
#include "occ_cache.hpp"

namespace ssb {

const char* config_name(int cfg)
{
    switch (cfg) {
    case CFG_DIRECT_16: return "direct16";
    case CFG_DIRECT_32: return "direct32";
    case CFG_TILE_64: return "tile64";
    case CFG_TILE_128: return "tile128";
    case CFG_TILE_128x64: return "tile128x64";
    case CFG_TILE_64x128: return "tile64x128";
    case CFG_TILE_128_K2: return "tile128k2";
    case CFG_TILE_128_W8: return "tile128w8";
    case CFG_TILE_256x128: return "tile256x128";
    case CFG_TILE_64_S2: return "tile64s2";
    case CFG_TILE_128_S2: return "tile128s2";
    case CFG_TILE_128_K2_S2: return "tile128k2s2";
    case CFG_TILE_128_W8_S2: return "tile128w8s2";
    case CFG_TILE_256x128_S2: return "tile256x128s2";
    case CFG_DMA_64: return "dma64";
    case CFG_DMA_128: return "dma128";
    case CFG_DMA_128_K2: return "dma128k2";
    case CFG_DMA_256x128: return "dma256x128";
    case CFG_DMA_128_S3: return "dma128s3";
    case CFG_KC_256x128: return "kc256x128";
    case CFG_KC_128: return "kc128";
    case CFG_DMA_256: return "dma256";
    case CFG_DMA_256_W16: return "dma256w16";
    case CFG_DMA_256_K2: return "dma256k2";
    case CFG_GDMA_128: return "gdma128";
    case CFG_GDMA_64: return "gdma64";
    case CFG_GDMA_256x128: return "gdma256x128";
    default: return "unknown";
    }
}

void config_tile(int cfg, int* tm, int* tn)
{
    static const int dims[CFG_COUNT][2] = {{16, 16},   {32, 32},   {64, 64},   {128, 128}, {128, 64},
                                           {64, 128},  {128, 128}, {128, 128}, {256, 128}, {64, 64},
                                           {128, 128}, {128, 128}, {128, 128}, {256, 128}, {64, 64},
                                           {128, 128}, {128, 128}, {256, 128}, {128, 128},
                                           {256, 128}, {128, 128}, {256, 256}, {256, 256}, {256, 256},
                                           {128, 128}, {64, 64},   {256, 128}};
    if (cfg < 0 || cfg >= CFG_COUNT) cfg = CFG_TILE_64;
    *tm = dims[cfg][0];
    *tn = dims[cfg][1];
}

OccupancyCache& occupancy_cache()
{
    static OccupancyCache cache;
    return cache;
}

int device_cu_count()
{
    static int cached[16] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
    if (cached[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        cached[dev] = cus;
    }
    return cached[dev];
}

static long tiles_for(const GemmProblem& g, int cfg)
{
    int tm, tn;
    config_tile(cfg, &tm, &tn);
    return (long)g.batch * ((g.m + tm - 1) / tm) * ((g.n + tn - 1) / tn);
}

int default_config(const GemmProblem& g)
{
    if (g.m <= 16 && g.n <= 16) return CFG_DIRECT_16;
    if (g.m <= 64 && g.n <= 64) return CFG_DIRECT_32;
    // Large tiles when they still give every CU at least two tiles.
    long cus = device_cu_count();
    if (g.m >= 128 && g.n >= 128 && tiles_for(g, CFG_TILE_128) >= 2 * cus) return CFG_TILE_128;
    if (g.m >= 128 && g.n < 128 && tiles_for(g, CFG_TILE_128x64) >= 2 * cus) return CFG_TILE_128x64;
    if (g.m < 128 && g.n >= 128 && tiles_for(g, CFG_TILE_64x128) >= 2 * cus) return CFG_TILE_64x128;
    return CFG_TILE_64;
}

// SSBLAS_TILE_CONFIG=<name> forces a config for every call (tuning runs).
static int forced_config()
{
    const char* e = getenv("SSBLAS_TILE_CONFIG");
    if (e == nullptr) return -1;
    for (int c = 0; c < CFG_COUNT; ++c)
        if (strcmp(e, config_name(c)) == 0) return c;
    return -1;
}

int pick_config(const GemmProblem& g)
{
    int f = forced_config();
    if (f >= 0) return f;
    int t = tuned_config(g);
    if (t >= 0) return t;
    return default_config(g);
}

hipError_t launch_gemm(const GemmProblem& g, int cfg, hipStream_t stream)
{
    switch (g.arith) {
    case K_I8: return launch_i8(g, cfg, stream);
    case K_F32: return launch_f32(g, cfg, stream);
    case K_F64: return launch_f64(g, cfg, stream);
    case K_BF16X3: return launch_bf16x3(g, cfg, stream);
    case K_BF16X6: return launch_bf16x6(g, cfg, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ssb
