// grouped.hpp -- device data of a grouped plan (shared by grouped.hip and the
// host plan builder).  Plain structs, no device code.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "launch.hpp"

namespace ssb {

// Tile shapes of the grouped kernel (all 4 waves, LDS double-buffered).
enum GroupedCfg : int { GCFG_32 = 0, GCFG_64 = 1, GCFG_128 = 2, GCFG_COUNT = 3 };

struct GScalars {  // same layout as the device Scalars
    float af, bf;
    double ad, bd;
    int32_t ai, bi;
};

struct GGroup {
    int form;   // ArithKind
    int ta, tb;
    int beta0;
    GScalars s;
};

struct GInst {
    int m, n, k, group;
    long lda, ldb, ldc;
};

struct GItem {
    int inst;
    int i0, j0;
    int cfg;
};

struct GroupedDevice {
    const GItem* items;
    int nitems;
    const GInst* insts;
    const GGroup* groups;
    const void* const* A;
    const void* const* B;
    void* const* C;
};

hipError_t launch_grouped(const GroupedDevice& d, hipStream_t st);
int grouped_lds_bytes();

}  // namespace ssb
