// Host-only check of the occupancy cache (ssblasex_amd/csrc/occ_cache.hpp):
// two kernels of one function type keep separate entries, entries are per
// device and per block size, and each is queried once.  Exit status 0 = pass.
#include <stdio.h>

#include "occ_cache.hpp"

namespace ssb {
OccupancyCache& occupancy_cache()
{
    static OccupancyCache c;
    return c;
}
}  // namespace ssb

struct Params {
    int x;
};
// Two distinct functions of the same type, as the batched tile kernels are.
void kernel_a(Params) {}
void kernel_b(Params) {}

static int queries = 0;

static int occ(int dev, void (*k)(Params), int threads, int answer)
{
    return ssb::occupancy_cache().get(dev, reinterpret_cast<const void*>(k), threads, [&]() {
        ++queries;
        return answer;
    });
}

#define CHECK(c)                                          \
    do {                                                  \
        if (!(c)) {                                       \
            fprintf(stderr, "failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                     \
        }                                                 \
    } while (0)

int main()
{
    CHECK(occ(0, kernel_a, 512, 1) == 1);
    CHECK(occ(0, kernel_b, 256, 4) == 4);          // same type, own entry
    CHECK(occ(0, kernel_a, 512, 7) == 1);          // cached: not queried again
    CHECK(occ(0, kernel_b, 256, 9) == 4);
    CHECK(queries == 2);
    CHECK(occ(1, kernel_a, 512, 2) == 2);          // another device
    CHECK(occ(0, kernel_a, 256, 3) == 3);          // another block size
    CHECK(occ(0, kernel_b, 128, 0) == 1);          // a failed query counts as 1
    CHECK(queries == 5);
    CHECK(ssb::occupancy_cache().size() == 5);
    printf("ok\n");
    return 0;
}
