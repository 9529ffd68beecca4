// grouped_host.cpp -- grouped plans: argument checks, the tile shapes of
// each instance, the device list of entries in per-XCD queues with their
// claim counters, execution and graph capture.
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "api_internal.hpp"
#include "grouped.hpp"

struct ssblasGroupedPlan {
    int device = 0;
    void* mem = nullptr;  // entries and claim counters: one device allocation
    int nfam = 0;
    int fam_id[ssb::GF_COUNT] = {};
    ssb::GroupedDevice fam[ssb::GF_COUNT] = {};
    std::mutex mu;  // guards the lazily captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace ssb {

const char* grouped_cfg_name(int cfg)
{
    static const char* names[GCFG_COUNT] = {"g32", "g64", "g128", "w32"};
    return cfg >= 0 && cfg < GCFG_COUNT ? names[cfg] : "?";
}

hipError_t launch_grouped(int family, const GroupedDevice& d, hipStream_t st)
{
    switch (family) {
    case GF_I8_F32: return launch_grouped_i8_f32(d, st);
    case GF_I8_BF16X3: return launch_grouped_i8_bf16x3(d, st);
    case GF_I8_BF16X6: return launch_grouped_i8_bf16x6(d, st);
    default: return launch_grouped_f64(d, st);
    }
}

namespace {

int tile_of(int cfg) { return cfg == GCFG_32 || cfg == GCFG_W32 ? 32 : cfg == GCFG_64 ? 64 : 128; }

int elem_bytes_in(int form) { return form == K_I8 ? 1 : form == K_F64 ? 8 : 4; }
int elem_bytes_out(int form) { return form == K_F64 ? 8 : 4; }

// Estimated time (ns) of one item: the k-blocks at the slower of MFMA and
// memory rate for one of the two workgroups a CU holds, a floor per k-block
// for the load latency the double buffer does not cover, and a fixed cost
// for the first load and the epilogue.
double item_ns(int form, int cfg, long k, bool beta0)
{
    if (cfg == GCFG_W32) {
        // one wave of the four a workgroup runs at once: a quarter of the
        // workgroup tile's rate, no staging latency per k-block
        const double per_cu = 4.0 / 2.0;  // waves per CU running at once / workgroup tile ratio
        return (item_ns(form, GCFG_32, k, beta0) - 1000.0) * per_cu + 500.0;
    }
    // per-CU dense MFMA rate in ops/ns at 2.4 GHz, split forms at their product count
    static const double rate[5] = {8192 * 2.4, 256 * 2.4, 128 * 2.4, 4096 * 2.4 / 3, 4096 * 2.4 / 6};
    const double cu_bw = 8000.0 / 256.0;  // bytes/ns per CU with every CU busy
    const int t = tile_of(cfg), bk = grouped_bk(form, cfg);
    const long nk = (k + bk - 1) / bk;
    const double mma = 2.0 * t * t * bk / (rate[form] / 2);
    const double mem = 2.0 * t * bk * elem_bytes_in(form) / (cu_bw / 2);
    const double epi = (double)t * t * elem_bytes_out(form) * (beta0 ? 1 : 2) / (cu_bw / 2);
    return nk * std::max(std::max(mma, mem), 250.0) + epi + 1000.0;
}

long tiles(int cfg, long m, long n)
{
    const long t = tile_of(cfg);
    return ((m + t - 1) / t) * ((n + t - 1) / t);
}

// SSBLAS_GROUPED_CFG forces the shape (tuning and test runs): "<shape>" for
// every instance, or "<shape>:<shape>" for instances with m, n <= 32 and
// for the others.
int forced_cfg(long m, long n)
{
    const char* e = getenv("SSBLAS_GROUPED_CFG");
    if (e == nullptr) return -1;
    char name[16];
    const char* colon = strchr(e, ':');
    if (colon != nullptr) {
        const bool small = m <= 32 && n <= 32;
        const char* src = small ? e : colon + 1;
        const size_t len = small ? (size_t)(colon - e) : strlen(colon + 1);
        if (len >= sizeof(name)) return -1;
        memcpy(name, src, len);
        name[len] = 0;
    } else {
        if (strlen(e) >= sizeof(name)) return -1;
        strcpy(name, e);
    }
    for (int c = 0; c < GCFG_COUNT; ++c)
        if (strcmp(name, grouped_cfg_name(c)) == 0) return c;
    return -1;
}

// Shape of the strip that covers `r` rows (columns) left over past the last
// whole tile of shape `cfg`: the smallest shape at least r wide (one-wave
// 32x32 items for r <= 32, where the form has them), `cfg` itself when none
// is smaller.  int8 keeps whole tiles: its tiles are bound by memory, not by
// the MFMA work of the rows past the edge, and the one-wave items, which
// have one k-block in flight, would be slower.
int edge_cfg(int cfg, int form, int r)
{
    if (r <= 0 || form == K_I8) return cfg;
    const bool waves = form == K_I8 || form == K_F32;
    if (r <= 32 && tile_of(cfg) > 32) return waves ? GCFG_W32 : GCFG_32;
    if (r <= 64 && tile_of(cfg) > 64) return GCFG_64;
    return cfg;
}

// SSBLAS_GROUPED_EDGES=0 keeps whole tiles of one shape over every instance
// (tuning and test runs compare the two).
bool split_edges()
{
    const char* e = getenv("SSBLAS_GROUPED_EDGES");
    return e == nullptr || strcmp(e, "0") != 0;
}

int pick_cfg(int form, long m, long n, long k, bool beta0)
{
    int top = grouped_max_cfg(form);
    const int forced = forced_cfg(m, n);
    if (forced >= 0) return forced == GCFG_W32 ? forced : std::min(forced, top);
    int tuned = grouped_tuned_config(form, m, n, k);
    if (tuned >= 0) return tuned == GCFG_W32 ? tuned : std::min(tuned, top);
    int best = GCFG_32;
    double best_ns = 1e300;
    for (int c = GCFG_32; c <= top; ++c) {
        double ns = tiles(c, m, n) * item_ns(form, c, k, beta0);
        if (ns < best_ns) {
            best_ns = ns;
            best = c;
        }
    }
    if (tiles(GCFG_W32, m, n) * item_ns(form, GCFG_W32, k, beta0) < best_ns) best = GCFG_W32;
    return best;
}

long stored_rows(ssblasOperation_t op, long rows_of_op, long cols_of_op)
{
    return op == SSBLAS_OP_N ? rows_of_op : cols_of_op;
}

// Bytes a buffer descriptor for the stored matrix spans; 0 when it (with the
// read-ahead of one tile past its edge) does not fit 31-bit offsets.
uint32_t extent_bytes(long rows, long cols, long ld, int e)
{
    const long ext = ((cols - 1) * ld + rows) * (long)e;
    const long reach = ext + 256L * (ld + 1) * e;
    if (reach >= (1L << 31)) return 0;
    return (uint32_t)ext;
}

// What the plan keeps per group and per instance while it builds the records.
struct HostGroup {
    int form;   // ArithKind
    int akc;    // op(A) stored K-contiguous (A transposed)
    int bkc;    // op(B) stored K-contiguous (B not transposed)
    int beta0;
    uint32_t alpha[2], beta[2];  // scalar bits as in GRec
};
struct HostInst {
    int m, n, k, group;
    long lda, ldb, ldc;
    uint32_t abytes, bbytes;
};

struct HostItem {
    GRec rec;
    double ns;
};

// The group's alpha / beta in the bits the kernel reads for its form.
void scalar_bits(int form, const GemmProblem& sc, uint32_t (&a)[2], uint32_t (&b)[2])
{
    a[1] = b[1] = 0u;
    if (form == K_I8) {
        a[0] = (uint32_t)sc.ai;
        b[0] = (uint32_t)sc.bi;
    } else if (form == K_F64) {
        uint64_t x, y;
        memcpy(&x, &sc.ad, 8);
        memcpy(&y, &sc.bd, 8);
        a[0] = (uint32_t)x;
        a[1] = (uint32_t)(x >> 32);
        b[0] = (uint32_t)y;
        b[1] = (uint32_t)(y >> 32);
    } else {
        memcpy(&a[0], &sc.af, 4);
        memcpy(&b[0], &sc.bf, 4);
    }
}

}  // namespace

ssblasStatus_t grouped_create(ssblasGroupedPlan_t* plan, long group_count, const SSINT group_size[],
                              const ssblasOperation_t transa_array[], const ssblasOperation_t transb_array[],
                              const ssblasDataType_t Atype_array[], const ssblasDataType_t Btype_array[],
                              const ssblasDataType_t Ctype_array[], const void* const alpha_array[],
                              const void* const beta_array[], const SSINT m_array[], const SSINT n_array[],
                              const SSINT k_array[], const void* const A[], const SSINT lda_array[],
                              const void* const B[], const SSINT ldb_array[], void* const C[],
                              const SSINT ldc_array[], ssblasComputeType_t computeType)
{
    if (plan == nullptr) return SSBLAS_STATUS_INVALID_VALUE;
    *plan = nullptr;
    if (group_count < 0) return SSBLAS_STATUS_INVALID_VALUE;
    if (group_count == 0) return SSBLAS_STATUS_INTERNAL_ERROR;  // zero tasks, as batchCount == 0
    if (!group_size || !transa_array || !transb_array || !Atype_array || !Btype_array || !Ctype_array ||
        !alpha_array || !beta_array)
        return SSBLAS_STATUS_INVALID_VALUE;
    if (group_count > INT_MAX) return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;

    // computeType: applies to the fp32 groups (plain or a bf16 form) and to
    // fp64 groups (plain only); int8 groups ignore it.
    const int ct = (int)computeType;
    int fp32_form = K_F32;
    if (ct == SSBLAS_COMPUTE_32F_FAST_BF16X3) fp32_form = K_BF16X3;
    else if (ct == SSBLAS_COMPUTE_32F_EMULATED_BF16X6) fp32_form = K_BF16X6;
    else if (ct == SSBLAS_COMPUTE_64F_EMULATED_OZAKI ||
             (ct >= SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S3 && ct <= SSBLAS_COMPUTE_64F_EMULATED_OZAKI_S9))
        return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;  // grouped fp64 runs the plain form only
    else if (ct != SSBLAS_COMPUTE_DEFAULT_TYPE && ct != SSBLAS_COMPUTE_32I && ct != SSBLAS_COMPUTE_32F &&
             ct != SSBLAS_COMPUTE_64F)
        return SSBLAS_STATUS_INVALID_VALUE;

    long total = 0;
    for (long g = 0; g < group_count; ++g) {
        if (group_size[g] < 0) return SSBLAS_STATUS_INVALID_VALUE;
        total += group_size[g];
        if (total > INT_MAX) return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;
    }
    if (total == 0) return SSBLAS_STATUS_INTERNAL_ERROR;
    if (!m_array || !n_array || !k_array || !lda_array || !ldb_array || !ldc_array) return SSBLAS_STATUS_INVALID_VALUE;

    std::vector<HostGroup> groups((size_t)group_count);
    for (long g = 0; g < group_count; ++g) {
        TypeTriple tt = classify_types(Atype_array[g], Btype_array[g], Ctype_array[g]);
        if (tt == TT_NONE) return SSBLAS_STATUS_INVALID_VALUE;
        const ssblasOperation_t ta = transa_array[g], tb = transb_array[g];
        if ((ta != SSBLAS_OP_N && ta != SSBLAS_OP_T) || (tb != SSBLAS_OP_N && tb != SSBLAS_OP_T))
            return SSBLAS_STATUS_INVALID_VALUE;
        if (alpha_array[g] == nullptr || beta_array[g] == nullptr) return SSBLAS_STATUS_INVALID_VALUE;
        GemmProblem sc;
        fill_scalars(tt, alpha_array[g], beta_array[g], sc);
        HostGroup& G = groups[(size_t)g];
        G.form = tt == TT_I8 ? K_I8 : tt == TT_F64 ? K_F64 : fp32_form;
        G.akc = ta == SSBLAS_OP_T ? 1 : 0;
        G.bkc = tb == SSBLAS_OP_N ? 1 : 0;
        G.beta0 = sc.beta0 ? 1 : 0;
        scalar_bits(G.form, sc, G.alpha, G.beta);
    }

    std::vector<HostInst> insts((size_t)total);
    long idx = 0;
    for (long g = 0; g < group_count; ++g) {
        const ssblasOperation_t ta = transa_array[g], tb = transb_array[g];
        const int form = groups[(size_t)g].form;
        const int e = elem_bytes_in(form);
        for (long i = 0; i < group_size[g]; ++i, ++idx) {
            const long m = m_array[idx], n = n_array[idx], k = k_array[idx];
            const long lda = lda_array[idx], ldb = ldb_array[idx], ldc = ldc_array[idx];
            if (m < 0 || n < 0 || k < 0) return SSBLAS_STATUS_INVALID_VALUE;
            const long ra = stored_rows(ta, m, k), ca = stored_rows(ta, k, m);
            const long rb = stored_rows(tb, k, n), cb = stored_rows(tb, n, k);
            if (lda < std::max(ra, 1L) || ldb < std::max(rb, 1L) || ldc < std::max(m, 1L))
                return SSBLAS_STATUS_INVALID_VALUE;
            if (m > INT_MAX || n > INT_MAX || k > INT_MAX || ldc > INT_MAX) return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;
            HostInst& I = insts[(size_t)idx];
            I.m = (int)m;
            I.n = (int)n;
            I.k = (int)k;
            I.group = (int)g;
            I.lda = lda;
            I.ldb = ldb;
            I.ldc = ldc;
            I.abytes = I.bbytes = 0;
            if (m == 0 || n == 0 || k == 0) continue;  // nothing to compute (k == 0: C unchanged)
            I.abytes = extent_bytes(ra, ca, lda, e);
            I.bbytes = extent_bytes(rb, cb, ldb, e);
            if (I.abytes == 0 || I.bbytes == 0) return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;
        }
    }
    if (A == nullptr || B == nullptr || C == nullptr) return SSBLAS_STATUS_INVALID_VALUE;

    // Work lists per family, longest first.  Every tile of an instance, and
    // its one-wave items, go to one of the GQ XCD queues (instances longest
    // first onto the queue with the least estimated work so far).  A queue's
    // entries are its tiles and bundles of four one-wave items of similar
    // cost (one item per wave of the workgroup that claims the bundle), so
    // the small items run beside the tiles instead of after them.
    const int other_family = fp32_form == K_BF16X3 ? GF_I8_BF16X3 : fp32_form == K_BF16X6 ? GF_I8_BF16X6 : GF_I8_F32;
    struct InstWork {
        int f;
        double ns;
        size_t begin, end;  // its items in `items`
    };
    std::vector<HostItem> items;
    std::vector<InstWork> work;
    for (long i = 0; i < total; ++i) {
        const HostInst& I = insts[(size_t)i];
        if (I.m == 0 || I.n == 0 || I.k == 0) continue;
        const HostGroup& G = groups[(size_t)I.group];
        const int cfg = pick_cfg(G.form, I.m, I.n, I.k, G.beta0 != 0);
        const int f = G.form == K_F64 ? GF_F64 : other_family;
        const size_t begin = items.size();
        double inst_ns = 0.0;
        auto emit = [&](int c, int i0, int j0) {
            GRec r;
            r.inst = (int32_t)i;
            r.i0 = i0;
            r.j0 = j0;
            r.bits = grec_bits(c, G.form, G.akc, G.bkc, G.beta0);
            r.m = I.m;
            r.n = I.n;
            r.k = I.k;
            r.lda = (int32_t)I.lda;  // below 2^31: the extent fits 31 bits
            r.ldb = (int32_t)I.ldb;
            r.ldc = (int32_t)I.ldc;
            r.abytes = I.abytes;
            r.bbytes = I.bbytes;
            memcpy(r.alpha, G.alpha, sizeof(r.alpha));
            memcpy(r.beta, G.beta, sizeof(r.beta));
            const HostItem h{r, item_ns(G.form, c, I.k, G.beta0 != 0)};
            items.push_back(h);
            inst_ns += c == GCFG_W32 ? h.ns / 4 : h.ns;  // four one-wave items run at once
        };
        // Whole tiles of the chosen shape, then the rows (columns) left over
        // past the last whole tile in strips of the smallest shape that
        // covers them: a 128-tile over 22 leftover rows would spend its
        // whole k-loop on 106 rows that are never stored.
        const int t = tile_of(cfg);
        const int mf = split_edges() ? I.m / t * t : I.m, nf = split_edges() ? I.n / t * t : I.n;
        const int em = edge_cfg(cfg, G.form, I.m - mf), en = edge_cfg(cfg, G.form, I.n - nf);
        const int tm = tile_of(em), tn = tile_of(en);
        for (int j0 = 0; j0 < nf; j0 += t)
            for (int i0 = 0; i0 < mf; i0 += t) emit(cfg, i0, j0);
        if (mf < I.m)  // bottom strip: leftover rows under the whole tiles' columns
            for (int j0 = 0; j0 < nf; j0 += tm) emit(em, mf, j0);
        if (nf < I.n)  // right strip: leftover columns beside the whole tiles' rows
            for (int i0 = 0; i0 < mf; i0 += tn) emit(en, i0, nf);
        if (mf < I.m && nf < I.n) {  // corner
            const int ec = tile_of(em) >= tile_of(en) ? em : en;
            const int tc = tile_of(ec);
            for (int j0 = nf; j0 < I.n; j0 += tc)
                for (int i0 = mf; i0 < I.m; i0 += tc) emit(ec, i0, j0);
        }
        if (items.size() > begin) work.push_back(InstWork{f, inst_ns, begin, items.size()});
    }
    std::vector<size_t> order(work.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work[a].ns > work[b].ns; });
    double load[GF_COUNT][GQ] = {};
    std::vector<HostItem> qitems[GF_COUNT][GQ];
    for (size_t w : order) {
        const InstWork& W = work[w];
        int best = 0;
        for (int qi = 1; qi < GQ; ++qi)
            if (load[W.f][qi] < load[W.f][best]) best = qi;
        load[W.f][best] += W.ns;
        for (size_t x = W.begin; x < W.end; ++x) qitems[W.f][best].push_back(items[x]);
    }

    // Entries: a tile's record four times (each wave reads record 4 e + its
    // index), or a bundle of four one-wave items (a bundle short of four is
    // filled with empty items, m = n = k = 0).
    struct HostEntry {
        GRec rec[4];
        double ns;
    };
    std::vector<HostEntry> qlists[GF_COUNT][GQ];
    long nitems = 0;
    for (int f = 0; f < GF_COUNT; ++f) {
        for (int qi = 0; qi < GQ; ++qi) {
            // The queue's items are instance by instance, longest instance
            // first, each instance's tiles before its strips: kept in that
            // order, so that the entries of one instance run close together
            // in time and its operands are read from HBM into the XCD's L2
            // once.  One sequence per form (int8, the others), bundles of
            // four consecutive one-wave items of a form.
            std::vector<HostEntry> part[2];
            double tot[2] = {0.0, 0.0};
            HostEntry bundle[2];
            int nb[2] = {0, 0};
            auto close_bundle = [&](int x) {
                for (int w = nb[x]; w < 4; ++w) {
                    bundle[x].rec[w] = bundle[x].rec[0];
                    bundle[x].rec[w].m = bundle[x].rec[w].n = bundle[x].rec[w].k = 0;
                }
                part[x].push_back(bundle[x]);
                tot[x] += bundle[x].ns;
                nb[x] = 0;
            };
            for (const HostItem& h : qitems[f][qi]) {
                const int x = (int)((h.rec.bits >> 2) & 7u) == K_I8 ? 1 : 0;
                if ((h.rec.bits & 3u) == (uint32_t)GCFG_W32) {
                    if (nb[x] == 0) bundle[x].ns = 0.0;
                    bundle[x].rec[nb[x]++] = h.rec;
                    bundle[x].ns = std::max(bundle[x].ns, h.ns);
                    if (nb[x] == 4) close_bundle(x);
                } else {
                    HostEntry e;
                    for (int w = 0; w < 4; ++w) e.rec[w] = h.rec;
                    e.ns = h.ns;
                    part[x].push_back(e);
                    tot[x] += e.ns;
                }
            }
            for (int x = 0; x < 2; ++x)
                if (nb[x] > 0) close_bundle(x);
            // the forms merged so that each has done the same fraction of its
            // estimated work at every point: the int8 entries (bound by
            // memory) run beside the fp32 ones (bound by the MFMA pipes)
            // instead of after them
            std::vector<HostEntry>& E = qlists[f][qi];
            size_t at[2] = {0, 0};
            double done[2] = {0.0, 0.0};
            while (at[0] < part[0].size() || at[1] < part[1].size()) {
                int x;
                if (at[0] >= part[0].size()) x = 1;
                else if (at[1] >= part[1].size()) x = 0;
                else x = (done[0] + part[0][at[0]].ns) / tot[0] <= (done[1] + part[1][at[1]].ns) / tot[1] ? 0 : 1;
                done[x] += part[x][at[x]].ns;
                E.push_back(part[x][at[x]++]);
            }
            nitems += (long)E.size();
        }
    }
    if (nitems > INT_MAX) return SSBLAS_STATUS_NOTIMPLEMENTED_ERROR;

    // Device image: the records of each family
    const size_t off_items = 0;
    // then each family's claim counters (zero: the kernel re-arms them itself)
    const size_t off_ctrs = (size_t)std::max(nitems, 1L) * 4 * sizeof(GRec);
    constexpr size_t CTR_BYTES = (size_t)GCTR_COUNT * GCTR_STRIDE * sizeof(int);
    const size_t bytes = off_ctrs + GF_COUNT * CTR_BYTES;
    // the plan's device image is library-owned workspace
    if (bytes > workspace_limit()) return SSBLAS_STATUS_ALLOC_FAILED;
    if (!device_accessible(A) || !device_accessible(B) || !device_accessible(C))
        return SSBLAS_STATUS_INVALID_VALUE;
    std::vector<char> image(bytes, 0);

    ssblasGroupedPlan* p = new (std::nothrow) ssblasGroupedPlan();
    if (p == nullptr) return SSBLAS_STATUS_ALLOC_FAILED;
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipMalloc(&p->mem, bytes);
    if (e != hipSuccess) {
        delete p;
        return hip_status(e);
    }
    char* dev = static_cast<char*>(p->mem);
    size_t pos = off_items;
    for (int f = 0; f < GF_COUNT; ++f) {
        size_t nq = 0;
        for (int qi = 0; qi < GQ; ++qi) nq += qlists[f][qi].size();
        if (nq == 0) continue;
        GRec* dst = reinterpret_cast<GRec*>(image.data() + pos);
        GroupedDevice& D = p->fam[p->nfam];
        size_t x = 0;
        for (int qi = 0; qi < GQ; ++qi) {
            D.qoff[qi] = (int)x;
            for (const HostEntry& h : qlists[f][qi]) {
                for (int w = 0; w < 4; ++w) dst[4 * x + (size_t)w] = h.rec[w];
                ++x;
            }
        }
        D.qoff[GQ] = (int)x;
        D.recs = reinterpret_cast<const GRec*>(dev + pos);
        D.nitems = (int)nq;
        D.A = A;
        D.B = B;
        D.C = C;
        D.ctrs = reinterpret_cast<int*>(dev + off_ctrs + (size_t)p->nfam * CTR_BYTES);
        p->fam_id[p->nfam] = f;
        ++p->nfam;
        pos += nq * 4 * sizeof(GRec);
    }
    e = hipMemcpy(p->mem, image.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p->mem);
        delete p;
        return hip_status(e);
    }
    *plan = p;
    return SSBLAS_STATUS_SUCCESS;
}

ssblasStatus_t grouped_execute(ssblasGroupedPlan_t p, hipStream_t stream)
{
    if (p == nullptr) return SSBLAS_STATUS_INVALID_VALUE;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->device) return SSBLAS_STATUS_INVALID_VALUE;
    for (int i = 0; i < p->nfam; ++i) {
        hipError_t e = launch_grouped(p->fam_id[i], p->fam[i], stream);
        if (e != hipSuccess) return hip_status(e);
    }
    return SSBLAS_STATUS_SUCCESS;
}

ssblasStatus_t grouped_graph_launch(ssblasGroupedPlan_t p, hipStream_t stream)
{
    if (p == nullptr) return SSBLAS_STATUS_INVALID_VALUE;
    {
        std::lock_guard<std::mutex> lock(p->mu);
        if (p->exec == nullptr) {
            hipStream_t cs = nullptr;
            hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
            if (e != hipSuccess) return hip_status(e);
            e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
            ssblasStatus_t st = SSBLAS_STATUS_SUCCESS;
            if (e == hipSuccess) {
                st = grouped_execute(p, cs);
                hipGraph_t g = nullptr;
                hipError_t e2 = hipStreamEndCapture(cs, &g);
                if (e2 == hipSuccess) p->graph = g;
                if (e == hipSuccess) e = e2;
            }
            if (e == hipSuccess && st == SSBLAS_STATUS_SUCCESS)
                e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
            (void)hipStreamDestroy(cs);
            if (st != SSBLAS_STATUS_SUCCESS) return st;
            if (e != hipSuccess) return hip_status(e);
        }
    }
    return hip_status(hipGraphLaunch(p->exec, stream));
}

ssblasStatus_t grouped_destroy(ssblasGroupedPlan_t p)
{
    if (p == nullptr) return SSBLAS_STATUS_SUCCESS;
    // work of this plan may still be queued on any stream
    (void)hipDeviceSynchronize();
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->mem) (void)hipFree(p->mem);
    delete p;
    return SSBLAS_STATUS_SUCCESS;
}

}  // namespace ssb

extern "C" {

ssblasStatus_t ssblasGroupedPlanCreate(ssblasGroupedPlan_t* plan, SSINT group_count, const SSINT group_size[],
                                       const ssblasOperation_t transa_array[],
                                       const ssblasOperation_t transb_array[], const ssblasDataType_t Atype_array[],
                                       const ssblasDataType_t Btype_array[], const ssblasDataType_t Ctype_array[],
                                       const void* const alpha_array[], const void* const beta_array[],
                                       const SSINT m_array[], const SSINT n_array[], const SSINT k_array[],
                                       const void* const A[], const SSINT lda_array[], const void* const B[],
                                       const SSINT ldb_array[], void* const C[], const SSINT ldc_array[],
                                       ssblasComputeType_t computeType, ssblasGemmAlgo_t)
{
    return ssb::grouped_create(plan, group_count, group_size, transa_array, transb_array, Atype_array, Btype_array,
                               Ctype_array, alpha_array, beta_array, m_array, n_array, k_array, A, lda_array, B,
                               ldb_array, C, ldc_array, computeType);
}

ssblasStatus_t ssblasGroupedPlanExecute(ssblasGroupedPlan_t plan, ssblasStream_t stream)
{
    return ssb::grouped_execute(plan, reinterpret_cast<hipStream_t>(stream));
}

ssblasStatus_t ssblasGroupedPlanGraphLaunch(ssblasGroupedPlan_t plan, ssblasStream_t stream)
{
    return ssb::grouped_graph_launch(plan, reinterpret_cast<hipStream_t>(stream));
}

ssblasStatus_t ssblasGroupedPlanDestroy(ssblasGroupedPlan_t plan) { return ssb::grouped_destroy(plan); }

ssblasStatus_t ssblasGemmGroupedBatchedEx(SSINT group_count, const SSINT group_size[],
                                          const ssblasOperation_t transa_array[],
                                          const ssblasOperation_t transb_array[],
                                          const ssblasDataType_t Atype_array[], const ssblasDataType_t Btype_array[],
                                          const ssblasDataType_t Ctype_array[], const void* const alpha_array[],
                                          const void* const beta_array[], const SSINT m_array[],
                                          const SSINT n_array[], const SSINT k_array[], const void* const A[],
                                          const SSINT lda_array[], const void* const B[], const SSINT ldb_array[],
                                          void* const C[], const SSINT ldc_array[], ssblasComputeType_t computeType,
                                          ssblasGemmAlgo_t algo)
{
    ssblasGroupedPlan_t plan = nullptr;
    ssblasStatus_t st = ssblasGroupedPlanCreate(&plan, group_count, group_size, transa_array, transb_array,
                                                Atype_array, Btype_array, Ctype_array, alpha_array, beta_array,
                                                m_array, n_array, k_array, A, lda_array, B, ldb_array, C, ldc_array,
                                                computeType, algo);
    if (st != SSBLAS_STATUS_SUCCESS) return st;
    st = ssb::grouped_execute(plan, nullptr);
    if (st == SSBLAS_STATUS_SUCCESS) st = ssb::hip_status(hipStreamSynchronize(nullptr));
    ssb::grouped_destroy(plan);
    return st;
}

}  // extern "C"
