// occupancy.hpp -- kernel_occupancy(): resident workgroups per CU of a
// kernel on the current device, through the per-kernel cache of
// occ_cache.hpp.  Device-side translation units only.
#pragma once
#include <hip/hip_runtime.h>

#include "occ_cache.hpp"

namespace ssb {

// Resident workgroups per CU of `kernel` at `threads` per workgroup on the
// current device (cached per kernel address, see occ_cache.hpp).
template <class K>
int kernel_occupancy(K kernel, int threads)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    return occupancy_cache().get(dev, reinterpret_cast<const void*>(kernel), threads, [&]() {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess) nb = 1;
        return nb;
    });
}

}  // namespace ssb
