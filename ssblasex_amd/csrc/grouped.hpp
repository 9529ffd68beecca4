// grouped.hpp -- device data of a grouped plan (shared by the grouped kernel
// translation units and the host plan builder).  Plain structs, no device
// code.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "launch.hpp"

namespace ssb {

// Tile shapes of the grouped kernel.  32, 64, 128: workgroup tiles (4 waves
// in a 2x2 grid, LDS double-buffered, operand layouts read per item at run
// time).  W32: one wave per 32x32 block, operands loaded straight into MFMA
// registers (no LDS, no barriers); each wave of a workgroup takes its own.
enum GroupedCfg : int { GCFG_32 = 0, GCFG_64 = 1, GCFG_128 = 2, GCFG_W32 = 3, GCFG_COUNT = 4 };

// Kernel families: the arithmetic forms one grouped kernel can run.  int8
// shares a kernel with whichever fp32 form the plan uses, so the usual
// mixed int8/fp32 plan is a single launch.  fp64 groups run in a kernel of
// their own.
enum GroupedFamily : int {
    GF_I8_F32 = 0,
    GF_I8_BF16X3 = 1,
    GF_I8_BF16X6 = 2,
    GF_F64 = 3,
    GF_COUNT = 4
};

// One work item: one output tile of one instance, as a 64-byte record that
// holds everything the kernel needs except the operand pointers (those are
// read from the caller's arrays when the item runs, so a plan follows
// changes of the pointers).  One record is one 16-lane vector load.
struct GRec {
    int32_t inst;      // instance index (into the caller's A / B / C arrays)
    int32_t i0, j0;    // tile origin in C
    uint32_t bits;     // cfg | form << 2 | akc << 5 | bkc << 6 | beta0 << 7
    int32_t m, n, k;
    int32_t lda, ldb, ldc;
    uint32_t abytes, bbytes;  // extent of the stored A / B in bytes
    uint32_t alpha[2];  // the group's scalars in their scale type's bits
    uint32_t beta[2];   // (fp32 and int32 in word 0, fp64 in both words)
};
static_assert(sizeof(GRec) == 64, "one record = 16 dwords");
constexpr int GREC_WORDS = 16;

// Fields of GRec::bits.
constexpr uint32_t grec_bits(int cfg, int form, int akc, int bkc, int beta0)
{
    return (uint32_t)cfg | ((uint32_t)form << 2) | ((uint32_t)akc << 5) | ((uint32_t)bkc << 6) |
           ((uint32_t)beta0 << 7);
}

// Workgroup items are split into GQ queues, one per XCD: every tile of an
// instance is in one queue, so the instance's operands are fetched into one
// XCD's L2 (workgroups are dispatched to the eight XCDs round-robin, so
// workgroup b runs on XCD b % 8 and claims from queue b % GQ first).
constexpr int GQ = 8;

// Claim counters of a family's launch (GroupedDevice::ctrs), each in its
// own 256-byte line: one per queue, one counting the workgroups that have
// finished.  Zero when a launch starts; the last workgroup to finish zeroes
// them again.
constexpr int GCTR_STRIDE = 64;  // ints between counters
constexpr int GCTR_DONE = GQ, GCTR_COUNT = GQ + 1;

struct GroupedDevice {
    // Entries queue by queue, each queue sorted by estimated cost, longest
    // first.  Entry e is records 4 e .. 4 e + 3, one per wave of the
    // workgroup that takes it: a tile (four copies of its record) or a
    // bundle of four one-wave items (empty ones have m = n = k = 0).
    const GRec* recs;
    int nitems;          // entries (all queues)
    int qoff[GQ + 1];    // queue q holds entries qoff[q] .. qoff[q + 1] - 1
    const void* const* A;  // the caller's device arrays of device pointers
    const void* const* B;
    void* const* C;
    int* ctrs;           // claim counters, GCTR_STRIDE ints apart
};

// k-depth of a grouped tile per form and shape.
constexpr int grouped_bk(int form, int cfg)
{
    return form == K_I8 ? 64 : form == K_F64 ? 16 : 32;
}
// Largest tile shape per form: the bf16 split forms keep two or three
// operand planes in LDS and fp64 holds twice the accumulator registers, so
// their 128x128 tile would hold the family's kernel to one workgroup per CU.
constexpr int grouped_max_cfg(int form)
{
    return (form == K_BF16X3 || form == K_BF16X6 || form == K_F64) ? GCFG_64 : GCFG_128;
}

// Names of the grouped tile shapes in the tuning table ("g32", "g64", "g128", "w32").
const char* grouped_cfg_name(int cfg);
// Work-class column value that marks grouped rows in the tuning table.
constexpr int GROUPED_WORK_CLASS = 9;
// Size class of one dimension for grouped tuning: <= 32, <= 64, <= 128, more.
int grouped_size_class(long x);
// Tuned grouped tile shape for an instance of the form, -1 when none.
int grouped_tuned_config(int form, long m, long n, long k);

// Enqueue one persistent launch of the family's kernel over d's items.
hipError_t launch_grouped(int family, const GroupedDevice& d, hipStream_t st);
hipError_t launch_grouped_i8_f32(const GroupedDevice& d, hipStream_t st);
hipError_t launch_grouped_i8_bf16x3(const GroupedDevice& d, hipStream_t st);
hipError_t launch_grouped_i8_bf16x6(const GroupedDevice& d, hipStream_t st);
hipError_t launch_grouped_f64(const GroupedDevice& d, hipStream_t st);

}  // namespace ssb
