// tune.cpp -- per-size-bucket tile-config table.
#include "api_internal.hpp"

namespace ssb {
int tuned_config(const GemmProblem&) { return -1; }
}  // namespace ssb
