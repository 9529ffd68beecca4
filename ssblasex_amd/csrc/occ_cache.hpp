// occ_cache.hpp -- resident workgroups per CU, cached per (device, kernel,
// block size).
//
// The persistent grids are sized from the occupancy of the kernel actually
// launched.  Many kernels share one C++ function type (every batched tile
// kernel is void(KParams)), so the cache is keyed by the kernel's address,
// not by its type: a cache keyed by type would hand the first kernel's
// occupancy to every later kernel of that type.
//
// Host-only C++ (no HIP types): the query is a callable, so the cache logic
// builds and is tested with the host compiler alone.
#pragma once
#include <map>
#include <mutex>
#include <tuple>

namespace ssb {

class OccupancyCache {
public:
    // Cached value for (dev, kernel, threads); on a miss, query() is called
    // once and its result (at least 1) is kept.
    template <class Query>
    int get(int dev, const void* kernel, int threads, Query query)
    {
        const Key key(dev, kernel, threads);
        std::lock_guard<std::mutex> lk(mu_);
        auto it = map_.find(key);
        if (it != map_.end()) return it->second;
        int nb = query();
        if (nb < 1) nb = 1;
        map_.emplace(key, nb);
        return nb;
    }
    size_t size()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return map_.size();
    }

private:
    using Key = std::tuple<int, const void*, int>;
    std::mutex mu_;
    std::map<Key, int> map_;
};

// The process-wide cache (defined in dispatch.cpp).
OccupancyCache& occupancy_cache();

}  // namespace ssb
