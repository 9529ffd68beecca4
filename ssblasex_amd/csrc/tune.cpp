// tune.cpp -- the per-size-bucket tile-config table written by
// `python -m ssblasex_amd.tune` and read here at the first call.
//
// Table lines: "form ta tb mb nb kb wb config us" (see tune.py for the bucket
// definitions).  Grouped-plan rows carry wb = GROUPED_WORK_CLASS, ta = tb = 0,
// size classes (<= 32, <= 64, <= 128, larger) in mb/nb/kb and a grouped tile
// name (g32, g64, g128).  Path: $SSBLAS_TUNE_TABLE, else tune/<arch>.txt next to the
// lib/ directory that holds this library.
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "api_internal.hpp"
#include "grouped.hpp"

namespace ssb {

namespace {
using Key = std::tuple<int, int, int, int, int, int, int>;
std::once_flag g_once;
std::map<Key, int> g_table;

int bucket(long x)
{
    int b = 0;
    while ((1L << b) < x) ++b;
    return b < 3 ? 3 : (b > 13 ? 13 : b);
}

int work_class(long m, long n, long batch)
{
    long tiles = batch * ((m + 127) / 128) * ((n + 127) / 128);
    return tiles < 256 ? 0 : (tiles < 2048 ? 1 : 2);
}

std::string default_path()
{
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&default_path), &info) == 0 || info.dli_fname == nullptr) return "";
    std::string so = info.dli_fname;
    size_t slash = so.rfind('/');
    std::string libdir = slash == std::string::npos ? "." : so.substr(0, slash);
    char arch[64] = "gfx950";
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        strncpy(arch, prop.gcnArchName, sizeof(arch) - 1);
        arch[sizeof(arch) - 1] = 0;
        char* colon = strchr(arch, ':');
        if (colon) *colon = 0;
    }
    return libdir + "/../tune/" + arch + ".txt";
}

void load_table()
{
    const char* env = getenv("SSBLAS_TUNE_TABLE");
    std::string path = env ? env : default_path();
    if (path.empty()) return;
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#') continue;
        int form, ta, tb, mb, nb, kb, wb;
        char name[64];
        if (sscanf(line, "%d %d %d %d %d %d %d %63s", &form, &ta, &tb, &mb, &nb, &kb, &wb, name) != 8) continue;
        for (int c = 0; c < CFG_COUNT; ++c)
            if (strcmp(name, config_name(c)) == 0) g_table[Key(form, ta, tb, mb, nb, kb, wb)] = c;
        for (int c = 0; c < GCFG_COUNT; ++c)
            if (strcmp(name, grouped_cfg_name(c)) == 0) g_table[Key(form, ta, tb, mb, nb, kb, wb)] = c;
    }
    fclose(f);
}
}  // namespace

int tuned_config(const GemmProblem& g)
{
    std::call_once(g_once, load_table);
    if (g_table.empty()) return -1;
    auto it = g_table.find(Key(g.arith, g.ta ? 1 : 0, g.tb ? 1 : 0, bucket(g.m), bucket(g.n), bucket(g.k),
                               work_class(g.m, g.n, g.batch)));
    return it == g_table.end() ? -1 : it->second;
}

int grouped_size_class(long x) { return x <= 32 ? 0 : x <= 64 ? 1 : x <= 128 ? 2 : 3; }

int grouped_tuned_config(int form, long m, long n, long k)
{
    std::call_once(g_once, load_table);
    if (g_table.empty()) return -1;
    auto it = g_table.find(Key(form, 0, 0, grouped_size_class(m), grouped_size_class(n), grouped_size_class(k),
                               GROUPED_WORK_CLASS));
    return it == g_table.end() ? -1 : it->second;
}

}  // namespace ssb
