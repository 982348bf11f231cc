// runtime.hip -- the C ABI of include/bpx.h: contexts, the resident residue
// arena, query sets (one query or a batch packed into scan groups), scans and
// their results.  One HIP stream per context; every device buffer is owned
// by a handle.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "bpx.h"
#include "kernels.hpp"
#include "table.hpp"

namespace bpx {
const int8_t* blosum62();
hipError_t launch_features(const uint8_t* residues, const uint64_t* soff, const uint32_t* slen, uint32_t n,
                           const uint16_t* word_counts, int word, float* out, int num_cus, hipStream_t stream);
}

namespace {

// the last failure's message of each thread (a fixed buffer: reporting a
// failure allocates nothing)
thread_local char g_err[512];

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                                         \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return fail(BPX_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

constexpr uint32_t kArenaAlign = 128;  // subject starts in the arena: whole cache lines per subject
constexpr uint32_t kArenaTail = 256;   // readable zero bytes after the last subject
constexpr int kMaxClasses = 128;       // launches per scan: (table placement, multi, waves) bins x slot width (<= 2 x 2 x 13 x 2)
constexpr uint32_t kMetaWords = 8;     // u64 words before the per-query stats
constexpr uint32_t kSegFloor = 1024;   // smallest per-wave record segment (64 B per record slot)
constexpr uint32_t kFlatChunk = 16;    // most subjects a wavefront claims at a time

// Size-keyed cache of device (or pinned host) allocations, so repeated scans
// do not call hipMalloc/hipFree.  Host pools pass their hipHostMalloc flags.
struct Pool {
    Pool() = default;
    explicit Pool(unsigned host_flags) : host_(true), flags_(host_flags) {}
    bool host_ = false;
    unsigned flags_ = 0;
    // tests: this many allocations first make a request HIP must refuse (a
    // real hipErrorOutOfMemory), so the release-and-retry path below runs
    int fail_first = 0;
    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> sizes_;
    ~Pool() {
        for (auto& kv : free_) release(kv.second);
    }
    void release(void* p) {
        if (host_)
            (void)hipHostFree(p);
        else
            (void)hipFree(p);
    }
    // frees every cached (unused) allocation
    void trim() {
        for (auto& kv : free_) {
            release(kv.second);
            sizes_.erase(kv.second);
        }
        free_.clear();
    }
    hipError_t alloc(void** p, size_t bytes) {
        return host_ ? hipHostMalloc(p, bytes, flags_) : hipMalloc(p, bytes);
    }
    void* get(size_t bytes) {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            free_.erase(it);
            return p;
        }
        void* p = nullptr;
        hipError_t e;
        if (fail_first > 0) {
            --fail_first;
            e = alloc(&p, size_t(1) << 62);
        } else {
            e = alloc(&p, bytes);
        }
        if (e != hipSuccess) {
            // the failure is handled here (the cache is released and the
            // allocation retried): clear it from the thread's last error,
            // which every kernel launch check reads
            (void)hipGetLastError();
            trim();
            if (alloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        sizes_[p] = bytes;
        return p;
    }
    void put(void* p) {
        if (!p) return;
        auto it = sizes_.find(p);
        if (it == sizes_.end()) {
            release(p);
            return;
        }
        free_.emplace(it->second, p);
    }
};

int32_t next_pow2(int64_t v) {
    int64_t n = 1;
    while (n < v) n <<= 1;
    return (int32_t)n;
}

}  // namespace

struct bpx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    int64_t lds_limit = 0;
    int64_t hbm_bytes = 0;
    hipStream_t lanes[2] = {nullptr, nullptr};  // scans alternate between two streams, so
    uint32_t next_lane = 0;                     // consecutive scans may overlap on the device
    uint32_t seg_hint = 0;             // per-wave record segment size a previous scan needed
    Pool pool;
    Pool pinned{hipHostMallocDefault};
    // counters the scan kernel writes to directly: coherent, so its stores
    // go to host memory as they are made, not to the device caches
    Pool coherent{hipHostMallocMapped | hipHostMallocCoherent};
    // this device's copy of the allocation-free scan's scratch array
    // (hipGetSymbolAddress at context creation), or null with the reason
    uint8_t* scratch = nullptr;
    char scratch_err[256] = {0};
};

struct bpx_db {
    bpx_ctx* ctx = nullptr;
    uint32_t n = 0;
    uint64_t nres = 0;
    uint64_t arena_bytes = 0;
    uint8_t* d_arena = nullptr;
    uint64_t* d_off = nullptr;
    uint32_t* d_len = nullptr;
    uint32_t* d_order = nullptr;  // all subjects, longest first
    uint64_t* d_poff = nullptr;   // arena offset of d_order[k]
    uint32_t* d_plen = nullptr;   // length of d_order[k]
    std::vector<uint64_t> h_off;
    std::vector<uint32_t> h_len;
    std::vector<uint32_t> h_order;
    bool cost_order = false;  // h_order set by bpx_db_set_cost
    // cost bins (set with a cost order): bin b is processing positions
    // [bin_end[b-1], bin_end[b]), of mean predicted cost bin_weight[b] times
    // the database's mean
    std::vector<uint32_t> bin_end;
    std::vector<float> bin_weight;
};
constexpr int kCostBins = 8;  // bins of equal predicted work per cost order
constexpr uint32_t kCostLongPrefix = 15360;  // cost orders keep subjects this long first, longest first
constexpr int32_t kCostMaxWindow = bpx::kNarrowLimit - kCostLongPrefix;

// Queries staged on the device as scan groups.
struct QuerySet {
    bpx_ctx* ctx = nullptr;
    uint32_t nq = 0;
    int32_t word = 3, window = 40;
    uint32_t ngroups_pv = 0;
    std::vector<bpx::GroupInfo> groups;
    std::vector<bpx::QueryInfo> qinfo;
    std::vector<uint32_t> max_cell;  // per group
    uint64_t table_bytes = 0;
    bpx::GroupInfo* d_groups = nullptr;
    bpx::QueryInfo* d_qinfo = nullptr;
    uint32_t* d_tables = nullptr;
    uint8_t* d_qres = nullptr;
    uint8_t* d_qmap = nullptr;  // the query of each group position, laid out like d_qres
    int8_t* d_matrix = nullptr;

    void release() {
        (void)hipFree(d_groups);
        (void)hipFree(d_qinfo);
        (void)hipFree(d_tables);
        (void)hipFree(d_qres);
        (void)hipFree(d_qmap);
        (void)hipFree(d_matrix);
        d_groups = nullptr;
        d_qinfo = nullptr;
        d_tables = nullptr;
        d_qres = nullptr;
        d_qmap = nullptr;
        d_matrix = nullptr;
    }
};

struct bpx_query {
    QuerySet set;
    int32_t qlen = 0;
    bpx_word_params params{};
    int32_t npresent = 0;
    int64_t nentries = 0;
};

struct bpx_batch {
    QuerySet set;
};

struct bpx_result {
    bpx_ctx* ctx = nullptr;
    const QuerySet* set = nullptr;
    const bpx_db* db = nullptr;
    uint32_t first = 0, count = 0;
    uint32_t flags = 0;
    bpx_launch launch{};
    bool waited = false;
    int reruns = 0;
    // device
    uint32_t* d_cnt = nullptr;
    int32_t* d_sstats = nullptr;           // device address of h_sstats
    uint32_t* d_run_loc = nullptr;
    uint32_t* d_run_len = nullptr;
    bpx::SegRecord* d_seg = nullptr;
    uint32_t seg_cap = 0;
    uint64_t seg_total = 0;
    uint32_t* d_tile = nullptr;
    bpx_init_hsp* d_records = nullptr;
    uint32_t* d_subjects = nullptr;
    uint32_t* d_queries = nullptr;
    unsigned long long* d_meta = nullptr;  // [0] total, [1] overflow word, [8..] 5 stats per query
    unsigned long long* d_phase = nullptr; // BPX_OUT_PHASES: 8 words per wave
    uint32_t* d_order = nullptr;           // owned only for sub-range scans
    uint64_t* d_poff = nullptr;            // (sub-range) descriptors in processing order
    uint32_t* d_plen = nullptr;
    bpx::WorkItem* d_items = nullptr;
    // host (pinned)
    unsigned long long* h_meta = nullptr;
    bpx_init_hsp* h_records = nullptr;
    uint32_t* h_subjects = nullptr;
    uint32_t* h_queries = nullptr;
    bpx_subject_stats* h_sstats = nullptr;
    uint64_t total = 0;
    // timing
    hipStream_t stream = nullptr;    // the context lane this scan runs on
    uint32_t* d_work = nullptr;      // item counters, one per launch class, then one subject counter per item
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    // graph mode (BPX_OUT_GRAPH): the captured device work of the scan
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_stale = false;  // buffers were re-allocated by a rescan
    float kernel_ms = 0.f, device_ms = 0.f;
    // launch plan
    struct Class {
        bool tab_lds = false, multi = false, narrow = true;
        uint32_t flat = 0;  // subjects a wave claims at a time (0: work items)
        int waves = 0;
        bpx::LdsPlan plan{};
        uint32_t item_first = 0, nitems = 0;
        uint32_t diag_cap = 0;
    };
    std::vector<Class> classes;
    std::vector<bpx::WorkItem> items;
    uint32_t lds_need = 0;  // largest per-workgroup LDS plan, including one that did not fit
};

namespace {

// ---- query sets -------------------------------------------------------------

struct QuerySpec {
    const uint8_t* seq;
    int32_t len;
    int32_t xdrop, cutoff, diag_len;
};

int build_set(bpx_ctx* ctx, const std::vector<QuerySpec>& qs, const int8_t* M, int32_t word, int32_t threshold,
              int32_t window, int32_t group_residues, int32_t group_queries, int32_t group_diag_slots, QuerySet* out,
              std::vector<bpx::HostTable>* keep_tables = nullptr) {
    QuerySet& set = *out;
    set.ctx = ctx;
    set.nq = (uint32_t)qs.size();
    set.word = word;
    set.window = window;
    set.ngroups_pv = bpx::packed_presence_groups(word);
    // Pack short queries (ascending length) into groups; long ones alone.
    std::vector<uint32_t> order(qs.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return qs[a].len < qs[b].len; });
    std::vector<std::vector<uint32_t>> members;
    for (size_t k = 0; k < order.size();) {
        std::vector<uint32_t> m{order[k]};
        int64_t tot = qs[order[k]].len, dsum = qs[order[k]].diag_len;
        ++k;
        if (tot < group_residues)
            while (k < order.size() && (int32_t)m.size() < group_queries && tot + qs[order[k]].len <= group_residues &&
                   (group_diag_slots <= 0 || dsum + qs[order[k]].diag_len <= group_diag_slots)) {
                dsum += qs[order[k]].diag_len;
                tot += qs[order[k]].len;
                m.push_back(order[k]);
                ++k;
            }
        members.push_back(std::move(m));
    }
    std::vector<uint32_t> tables;
    std::vector<std::vector<uint32_t>> directs;
    std::vector<uint8_t> qres, qmap;
    for (const auto& m : members) {
        std::vector<bpx::QuerySpan> spans;
        bpx::GroupInfo gi{};
        gi.qinfo_first = (uint32_t)set.qinfo.size();
        gi.nq = (uint32_t)m.size();
        gi.qres_off = (uint32_t)qres.size();
        uint32_t qstart = 0, dbase = 0;
        for (uint32_t id : m) {
            const QuerySpec& q = qs[id];
            spans.push_back({q.seq, q.len});
            bpx::QueryInfo qi{};
            qi.qstart = qstart;
            qi.qlen = (uint32_t)q.len;
            qi.dbase = dbase;
            qi.dmask = (uint32_t)q.diag_len - 1;
            qi.xdrop = q.xdrop;
            qi.cutoff = q.cutoff;
            qi.qid = id;
            set.qinfo.push_back(qi);
            qres.insert(qres.end(), q.seq, q.seq + q.len);
            qmap.insert(qmap.end(), (size_t)q.len, (uint8_t)(set.qinfo.size() - 1 - gi.qinfo_first));
            qstart += (uint32_t)q.len;
            dbase += (uint32_t)q.diag_len;
        }
        gi.qres_len = qstart;
        gi.diag_slots = dbase;
        bpx::HostTable t;
        if (!bpx::build_table_multi(spans.data(), (int32_t)spans.size(), M, threshold, word, &t))
            return fail(BPX_E_ARG, "lookup table build failed (residue codes must be < 28, group < 65536 residues)");
        const std::vector<uint32_t> blob = t.pack();
        gi.table_word = (uint32_t)tables.size();
        gi.table_words = (uint32_t)blob.size();
        gi.npresent = t.npresent;
        gi.flags = t.narrow_starts() ? bpx::kGroupNarrowStarts : 0u;
        tables.insert(tables.end(), blob.begin(), blob.end());
        directs.push_back(t.direct_starts());
        set.max_cell.push_back(t.max_cell);
        set.groups.push_back(gi);
        if (keep_tables) keep_tables->push_back(std::move(t));
    }
    // each group's start offsets by word, after the packed tables (a
    // launch that reads its table from HBM looks words up through these)
    for (size_t gi = 0; gi < directs.size(); ++gi) {
        if (tables.size() + directs[gi].size() > 0xFFFFFFFFull)
            return fail(BPX_E_BUDGET, "query tables exceed 2^32 words; use fewer queries per batch");
        set.groups[gi].direct_word = (uint32_t)tables.size();
        tables.insert(tables.end(), directs[gi].begin(), directs[gi].end());
    }
    set.table_bytes = (uint64_t)tables.size() * 4;
    // zero words after the last table: the scan reads start entries at a
    // clamped rank without a branch and may touch one word past a group
    tables.insert(tables.end(), 16, 0u);
    qres.insert(qres.end(), 16, 0u);
    qmap.insert(qmap.end(), 16, 0u);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t ng = std::max<size_t>(set.groups.size(), 1), nqi = std::max<size_t>(set.qinfo.size(), 1);
    if (hipMalloc(&set.d_groups, ng * sizeof(bpx::GroupInfo)) != hipSuccess ||
        hipMalloc(&set.d_qinfo, nqi * sizeof(bpx::QueryInfo)) != hipSuccess ||
        hipMalloc(&set.d_tables, std::max<size_t>(tables.size(), 1) * 4) != hipSuccess ||
        hipMalloc(&set.d_qres, std::max<size_t>(qres.size(), 1)) != hipSuccess ||
        hipMalloc(&set.d_qmap, std::max<size_t>(qmap.size(), 1)) != hipSuccess ||
        hipMalloc(&set.d_matrix, 28 * 28) != hipSuccess) {
        set.release();
        return fail(BPX_E_NOMEM, "device allocation for query tables failed");
    }
    hipError_t e = hipSuccess;
    if (!set.groups.empty())
        e = hipMemcpy(set.d_groups, set.groups.data(), set.groups.size() * sizeof(bpx::GroupInfo),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess && !set.qinfo.empty())
        e = hipMemcpy(set.d_qinfo, set.qinfo.data(), set.qinfo.size() * sizeof(bpx::QueryInfo), hipMemcpyHostToDevice);
    if (e == hipSuccess && !tables.empty())
        e = hipMemcpy(set.d_tables, tables.data(), tables.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && !qres.empty()) e = hipMemcpy(set.d_qres, qres.data(), qres.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !qmap.empty()) e = hipMemcpy(set.d_qmap, qmap.data(), qmap.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(set.d_matrix, M, 28 * 28, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set.release();
        return fail(BPX_E_HIP, "query upload failed: %s", hipGetErrorString(e));
    }
    return BPX_OK;
}

// ---- launch planning ----------------------------------------------------------

bpx_launch resolve_launch(const bpx_launch* in) {
    bpx_launch l{};
    if (in) l = *in;
    if (l.groups_per_cu <= 0) l.groups_per_cu = 1;
    if (l.subject_buffer <= 0) l.subject_buffer = 2048;
    uint32_t sb = 1024;
    while ((int32_t)sb < l.subject_buffer) sb <<= 1;
    l.subject_buffer = (int32_t)sb;
    return l;
}

// Largest per-group LDS dimensions of groups that share a launch.
struct Dims {
    uint32_t table = 0, qres = 0, nq = 1, diag = 0;
    void add(const bpx::GroupInfo& g) {
        table = std::max(table, g.table_words * 4);
        qres = std::max(qres, g.qres_len);
        nq = std::max(nq, g.nq);
        diag = std::max(diag, g.diag_slots);
    }
};

// every wave count from 16 down to 8: a packed group's per-wave diagonal
// tables often fit 13-15 waves (measured 2 % faster on the 1,024-query batch
// than rounding down to 12)
constexpr int kAutoWaves[] = {16, 15, 14, 13, 12, 11, 10, 9, 8, 6, 4, 2, 1};
// A table goes to LDS only when that keeps at least this many waves (or as
// many as an HBM table): plain scans of 800-1,000-aa queries took 4.8-7.8 ms
// with LDS tables at 8-11 waves against 4.2-4.6 ms with HBM tables at 16.
constexpr int kLdsMinWaves = 12;

// Groups are binned (one launch per bin) by coarse wave counts {16, 12, 8,
// 4, 2, 1}; a bin then runs at the most waves in its range that all its
// members fit.  Binning by exact wave count made more launches, each with
// its own tail: 9 % slower on the skewed 500,000-subject batch.
int wave_bucket(int w) { return w >= 16 ? 16 : w >= 12 ? 12 : w >= 8 ? 8 : w >= 4 ? 4 : w >= 2 ? 2 : w; }
int bucket_top(int b) { return b >= 16 ? 16 : b == 12 ? 15 : b == 8 ? 11 : b == 4 ? 7 : b == 2 ? 3 : b; }

// The launch-shape decisions of a scan of one query set over `count`
// subjects, the first `nwide` of which need wide diagonal slots.  Holds
// scalars only, so the allocation-free scan plans with it too.
struct Planner {
    bpx_ctx* ctx = nullptr;
    bpx_launch L{};
    int user_waves = 0;          // 0: automatic
    uint32_t qdef = 0, qalt = 0;  // hit ring: preferred, and smallest that holds a batch plus the largest cell
    uint32_t sbuf = 0, limit = 0;
    bool any_narrow = false, any_wide = false;

    Planner(bpx_ctx* c, const QuerySet& set, uint32_t count, uint32_t nwide, const bpx_launch* user)
        : ctx(c), L(resolve_launch(user)) {
        user_waves = (!user || user->waves_per_group <= 0) ? 0
                                                           : std::min(bpx::kMaxWaves, std::max(1, user->waves_per_group));
        uint32_t maxcell = 0;
        for (uint32_t m : set.max_cell) maxcell = std::max(maxcell, m);
        // hit ring: the requested capacity, else 256 entries, shrunk to the
        // smallest that holds a batch plus the largest cell when LDS is short
        auto pow2_at_least = [](uint32_t v) {
            uint32_t n = 1;
            while (n < v) n <<= 1;
            return n;
        };
        const uint32_t qmin = pow2_at_least(std::max<uint32_t>(128u, maxcell + 64u));
        qdef = L.queue_capacity > 0 ? std::max(qmin, pow2_at_least((uint32_t)L.queue_capacity)) : std::max(qmin, 256u);
        qalt = L.queue_capacity > 0 ? qdef : qmin;
        sbuf = (uint32_t)L.subject_buffer;
        limit = (uint32_t)ctx->lds_limit;
        any_narrow = nwide < count;
        any_wide = nwide > 0;
    }
    const int* cands() const { return user_waves ? &user_waves : kAutoWaves; }
    int ncands() const { return user_waves ? 1 : (int)(sizeof kAutoWaves / sizeof kAutoWaves[0]); }
    int fewest() const { return cands()[ncands() - 1]; }

    // Per-wave buffers for `waves` waves: the first of (claim table, ring) =
    // (256, default), (256, small), (128, default), (128, small) that fits.
    bpx::LdsPlan plan_of(const Dims& a, bool tab_lds, int waves, bool narrow) const {
        const uint32_t claims[4] = {256u, 256u, 128u, 128u}, rings[4] = {qdef, qalt, qdef, qalt};
        bpx::LdsPlan pl{};
        for (int k = 0; k < 4; ++k) {
            pl = bpx::plan_lds(a.table, a.qres, a.nq, a.diag, rings[k], sbuf, claims[k], waves, tab_lds, narrow);
            if (pl.total <= limit) break;
        }
        return pl;
    }
    // the most waves (of the candidates, at most `below`) at which `a` fits
    // for every slot width this scan needs; 0 if none
    int best_waves(const Dims& a, bool tab_lds, int below) const {
        for (int k = 0; k < ncands(); ++k) {
            const int w = cands()[k];
            if (w > below) continue;
            if ((!any_narrow || plan_of(a, tab_lds, w, true).total <= limit) &&
                (!any_wide || plan_of(a, tab_lds, w, false).total <= limit))
                return w;
        }
        return 0;
    }
    // Where one group runs with the most waves: its table in LDS when that
    // keeps at least kLdsMinWaves waves (or as many as the HBM table would),
    // else its table in HBM.  waves = 0: it fits nowhere.
    void place(const bpx::GroupInfo& g, bool* lds, int* waves) const {
        Dims one;
        one.add(g);
        const int wl = (g.flags & bpx::kGroupNarrowStarts) ? best_waves(one, true, bpx::kMaxWaves) : 0;
        const int wh = best_waves(one, false, bpx::kMaxWaves);
        *lds = wl > 0 && (wl >= kLdsMinWaves || wl >= wh);
        *waves = *lds ? wl : wh;
    }
    int refuse(const Dims& a, bool tab_lds, int waves, uint32_t* lds_need) const {
        const bpx::LdsPlan pl = plan_of(a, tab_lds, waves, any_narrow);
        *lds_need = std::max(*lds_need, pl.total);
        return fail(BPX_E_BUDGET,
                    "LDS plan %u B exceeds %u B (diagonal slots %u, %d waves, subject buffer %u); "
                    "use fewer waves per group or a smaller subject buffer",
                    pl.total, limit, a.diag, waves, sbuf);
    }
    // A launch class over subjects [lo, hi) of the processing order; the
    // caller sets its items.
    bpx_result::Class make_class(const Dims& a, bool tab_lds, bool multi, bool narrow, int waves, uint32_t lo,
                                 uint32_t hi) const {
        bpx_result::Class c;
        c.tab_lds = tab_lds;
        c.multi = multi;
        c.narrow = narrow;
        c.waves = waves;
        c.plan = plan_of(a, tab_lds, waves, narrow);
        c.diag_cap = std::max<uint32_t>(a.diag, 1);
        if (L.subjects_per_item <= 0) {
            // one item per group; every workgroup works through them in order
            // and its waves claim subjects device-wide, a few at a time: by
            // default about 1/16 of a wave's share (at most kFlatChunk), so the longest
            // subjects, which go first, cannot pile up on a few waves, and the
            // counters see few enough atomics.  256 per-query scans of 500,000
            // subjects took 0.80 s at 8 at a time, 0.84 s at 16 and 1.35 s at
            // 32; on 2,000,000 subjects 8, 16 and 32 are equal and 1 is 2.3x
            // slower (one atomic per subject).
            const uint64_t wv = (uint64_t)ctx->num_cus * L.groups_per_cu * waves;
            const uint64_t k = L.claims_per_share > 0 ? (uint64_t)L.claims_per_share : 16u;
            c.flat = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kFlatChunk, (hi - lo) / (wv * k)));
        }
        return c;
    }
};

// Splits the set's groups into launch classes and lays out LDS for each.
// Subjects [0, nwide) of the processing order (the longest, length + window
// >= kNarrowLimit) go to wide-slot launches, the rest to narrow-slot ones.
int plan_scan(bpx_result* r, const QuerySet& set, uint32_t count, uint32_t nwide, const bpx_launch* user) {
    const Planner P(r->ctx, set, count, nwide, user);
    // cost bins apply to scans of the whole database in its own order
    const bpx_db* cb = (r->db && !r->db->bin_end.empty() && r->first == 0 && count == r->db->n) ? r->db : nullptr;
    struct Acc {
        std::vector<uint32_t> groups;
        Dims d;
    };
    // Bin the groups: each goes where it runs with the most waves.  Groups of
    // one bin share a launch.  Key: (table in LDS, multi-query, waves).
    std::map<std::tuple<int, int, int>, Acc> bins;
    for (uint32_t gi = 0; gi < set.groups.size(); ++gi) {
        const bpx::GroupInfo& g = set.groups[gi];
        bool lds = false;
        int w = 0;
        P.place(g, &lds, &w);
        if (w == 0) {
            Dims one;
            one.add(g);
            return P.refuse(one, false, P.fewest(), &r->lds_need);
        }
        Acc& a = bins[std::make_tuple(lds ? 1 : 0, g.nq > 1 ? 1 : 0, wave_bucket(w))];
        a.groups.push_back(gi);
        a.d.add(g);
    }
    r->classes.clear();
    r->items.clear();
    // By default each launch class has one work item per group, and its
    // subjects are claimed device-wide by the wavefronts (Planner::make_class):
    // no workgroup barrier between items, so no wave waits for the slowest
    // wave of an item.  Runs of 8 items per CU (barriers between them) took
    // 11.2 ms on the 2,000,000-subject scan against 10.6 ms; per-query scans
    // of a 500,000-subject database took 12 % longer, a 1,024-query batch 2 %.
    // An explicit subjects_per_item keeps runs of that many subjects, each
    // claimed by one workgroup.
    const uint32_t R = (uint32_t)std::max(1, P.L.subjects_per_item);
    // LDS bins first, then by waves, most first
    std::vector<std::pair<std::tuple<int, int, int>, Acc*>> order;
    for (auto& kv : bins) order.push_back({kv.first, &kv.second});
    std::sort(order.begin(), order.end(), [](const auto& x, const auto& y) {
        if (std::get<0>(x.first) != std::get<0>(y.first)) return std::get<0>(x.first) > std::get<0>(y.first);
        if (std::get<2>(x.first) != std::get<2>(y.first)) return std::get<2>(x.first) > std::get<2>(y.first);
        return std::get<1>(x.first) < std::get<1>(y.first);
    });
    for (const auto& ob : order) {
        const Acc& a = *ob.second;
        const bool tab_lds = std::get<0>(ob.first) != 0;
        if (a.d.diag > 65536u)
            return fail(BPX_E_BUDGET, "a scan group needs %u diagonal slots (at most 65536); use smaller query groups",
                        a.d.diag);
        // the bin's largest members together may need fewer waves than each alone
        const int waves = P.best_waves(a.d, tab_lds, bucket_top(std::get<2>(ob.first)));
        if (waves == 0) return P.refuse(a.d, tab_lds, std::get<2>(ob.first), &r->lds_need);
        for (int w = 0; w < 2; ++w) {
            const bool narrow = w == 1;
            const uint32_t lo = narrow ? nwide : 0u, hi = narrow ? count : nwide;
            if (lo >= hi) continue;
            if ((int)r->classes.size() >= kMaxClasses) return fail(BPX_E_BUDGET, "too many launch classes");
            bpx_result::Class c = P.make_class(a.d, tab_lds, std::get<1>(ob.first) != 0, narrow, waves, lo, hi);
            r->lds_need = std::max(r->lds_need, c.plan.total);
            c.item_first = (uint32_t)r->items.size();
            if (c.flat && cb) {
                // a scan of the whole database with cost bins: one item per
                // (group, bin), the claim chunk scaled by the bin's mean cost
                for (uint32_t gi : a.groups) {
                    uint32_t f = lo;
                    for (size_t b = 0; b < cb->bin_end.size() && f < hi; ++b) {
                        const uint32_t e = std::min(hi, cb->bin_end[b]);
                        if (e <= f) continue;
                        const float w = std::max(cb->bin_weight[b], 1e-6f);
                        const uint32_t chunk = (uint32_t)std::max<long>(
                            1, std::min<long>(kFlatChunk, std::lround((float)c.flat / w)));
                        r->items.push_back({gi, f, e - f, chunk});
                        f = e;
                    }
                    if (f < hi) r->items.push_back({gi, f, hi - f, 0});
                }
            } else if (c.flat) {
                for (uint32_t gi : a.groups) r->items.push_back({gi, lo, hi - lo, 0});
            } else {
                for (uint32_t gi : a.groups)
                    for (uint32_t f = lo; f < hi; f += R) r->items.push_back({gi, f, std::min(R, hi - f), 0});
            }
            c.nitems = (uint32_t)r->items.size() - c.item_first;
            r->classes.push_back(c);
        }
    }
    r->launch = P.L;
    return BPX_OK;
}

// ---- results --------------------------------------------------------------------

void result_release(bpx_result* r, bool keep_order) {
    auto& pool = r->ctx->pool;
    auto& pinned = r->ctx->pinned;
    pool.put(r->d_cnt);
    pool.put(r->d_run_loc);
    pool.put(r->d_run_len);
    pool.put(r->d_seg);
    pool.put(r->d_tile);
    pool.put(r->d_records);
    pool.put(r->d_subjects);
    pool.put(r->d_queries);
    pool.put(r->d_meta);
    pool.put(r->d_phase);
    r->d_phase = nullptr;
    pool.put(r->d_items);
    pool.put(r->d_work);
    r->d_work = nullptr;
    if (!keep_order) {
        pool.put(r->d_order);
        pool.put(r->d_poff);
        pool.put(r->d_plen);
        r->d_order = nullptr;
        r->d_poff = nullptr;
        r->d_plen = nullptr;
    }
    pinned.put(r->h_records);
    pinned.put(r->h_subjects);
    pinned.put(r->h_queries);
    r->ctx->coherent.put(r->h_sstats);
    pinned.put(r->h_meta);
    r->d_cnt = nullptr;
    r->d_sstats = nullptr;
    r->d_run_loc = nullptr;
    r->d_run_len = nullptr;
    r->d_seg = nullptr;
    r->d_tile = nullptr;
    r->d_records = nullptr;
    r->d_subjects = nullptr;
    r->d_queries = nullptr;
    r->d_meta = nullptr;
    r->d_items = nullptr;
    r->h_records = nullptr;
    r->h_subjects = nullptr;
    r->h_queries = nullptr;
    r->h_sstats = nullptr;
    r->h_meta = nullptr;
}

uint64_t total_waves(const bpx_result* r) {
    uint64_t waves = 0;
    for (const auto& c : r->classes) waves += (uint64_t)r->ctx->num_cus * r->launch.groups_per_cu * c.waves;
    return waves;
}

int alloc_buffers(bpx_result* r, uint32_t seg_cap) {
    auto& pool = r->ctx->pool;
    const QuerySet& set = *r->set;
    const uint64_t pairs = (uint64_t)set.nq * r->count;
    const uint64_t runs = (uint64_t)set.groups.size() * r->count;
    r->seg_cap = seg_cap;
    r->seg_total = total_waves(r) * seg_cap;
    if (r->seg_total > 0xFFFFFFFFull) return fail(BPX_E_BUDGET, "record segments exceed 2^32 records");
    const uint64_t ntiles = (pairs + bpx::kTile - 1) / bpx::kTile;
    r->d_cnt = (uint32_t*)pool.get(pairs * 4);
    if ((r->flags & BPX_OUT_SUBJECT_STATS) && pairs) {
        r->h_sstats = (bpx_subject_stats*)r->ctx->coherent.get(pairs * sizeof(bpx_subject_stats));
        void* dev = nullptr;
        if (!r->h_sstats || hipHostGetDevicePointer(&dev, r->h_sstats, 0) != hipSuccess || !dev)
            return fail(BPX_E_NOMEM, "pinned allocation for subject stats failed (%llu query-subject pairs)",
                        (unsigned long long)pairs);
        r->d_sstats = static_cast<int32_t*>(dev);
    }
    r->d_run_loc = (uint32_t*)pool.get(runs * 4);
    r->d_run_len = (uint32_t*)pool.get(runs * 4);
    r->d_seg = (bpx::SegRecord*)pool.get(r->seg_total * sizeof(bpx::SegRecord));
    r->d_tile = (uint32_t*)pool.get(std::max<uint64_t>(ntiles, 1) * 4);
    r->d_records = (bpx_init_hsp*)pool.get(r->seg_total * sizeof(bpx_init_hsp));
    r->d_subjects = (uint32_t*)pool.get(r->seg_total * 4);
    r->d_queries = (uint32_t*)pool.get(r->seg_total * 4);
    r->d_meta = (unsigned long long*)pool.get((kMetaWords + 5ull * set.nq) * 8);
    if ((r->flags & BPX_OUT_PHASES) &&
        !(r->d_phase = (unsigned long long*)pool.get(std::max<uint64_t>(total_waves(r), 1) * 64)))
        return fail(BPX_E_NOMEM, "allocation for phase times failed");
    r->d_items = (bpx::WorkItem*)pool.get(std::max<size_t>(r->items.size(), 1) * sizeof(bpx::WorkItem));
    r->d_work = (uint32_t*)pool.get((kMaxClasses + r->items.size()) * 4);
    r->h_meta = (unsigned long long*)r->ctx->pinned.get((kMetaWords + 5ull * set.nq) * 8);
    if (!r->d_cnt || !r->d_run_loc || !r->d_run_len ||
        !r->d_seg || !r->d_tile || !r->d_records || !r->d_subjects || !r->d_queries || !r->d_meta || !r->d_items ||
        !r->d_work ||
        !r->h_meta)
        return fail(BPX_E_NOMEM, "allocation for scan results failed (%llu query-subject pairs, %llu segment records)",
                    (unsigned long long)pairs, (unsigned long long)r->seg_total);
    return BPX_OK;
}

// Everything the device work of one scan reads and writes: the result's
// pooled buffers, or the allocation-free scan's scratch.
struct ScanJob {
    bpx_ctx* ctx = nullptr;
    const QuerySet* set = nullptr;
    const bpx_db* db = nullptr;
    uint32_t first = 0, count = 0;
    const bpx_result::Class* classes = nullptr;
    size_t nclasses = 0, nitems = 0;
    int groups_per_cu = 1;
    uint32_t sbuf = 0;  // subject buffer bytes per wave
    hipStream_t stream = nullptr;
    const uint32_t* order = nullptr;  // processing order and descriptors in that order
    const uint64_t* poff = nullptr;
    const uint32_t* plen = nullptr;
    uint32_t* cnt = nullptr;
    int32_t* sstats = nullptr;  // optional per-subject counters
    uint32_t* run_loc = nullptr;
    uint32_t* run_len = nullptr;
    bpx::SegRecord* seg = nullptr;
    uint32_t seg_cap = 0;
    uint32_t* tile = nullptr;
    bpx_init_hsp* records = nullptr;
    uint32_t* subjects = nullptr;
    uint32_t* queries = nullptr;  // optional
    uint64_t rec_cap = ~0ull;     // records / subjects / queries capacity
    unsigned long long* meta = nullptr;
    bpx::WorkItem* items = nullptr;  // on the device
    uint32_t* work = nullptr;
    unsigned long long* h_meta = nullptr;  // host copy of the meta words
    unsigned long long* phase = nullptr;   // optional per-wave phase times
    hipEvent_t ev_start = nullptr, ev_kernels = nullptr;  // timing, optional
    bool capturing = false;  // the stream is being captured: kernel attributes are set beforehand
};

// The scan's device work on the job's stream: counter resets (fill kernels,
// no memset nodes), every launch class, compaction and the copy of the meta
// words to host memory (the scan kernel writes the subject counters itself).
// The events bracket the word-finder launches when given.  The work items
// must be on the device already.
int enqueue_job(const ScanJob& j) {
    const QuerySet& set = *j.set;
    hipStream_t st = j.stream;
    // meta words, work counters, and the counts and run lengths, which the
    // kernel writes only for subjects with records: one launch
    bpx::ZeroRanges z{};
    z.dst[0] = reinterpret_cast<uint32_t*>(j.meta);
    z.n[0] = 2 * (kMetaWords + 5ull * set.nq);
    z.dst[1] = j.work;
    z.n[1] = kMaxClasses + j.nitems;
    z.dst[2] = j.cnt;
    z.n[2] = (uint64_t)set.nq * j.count;
    z.dst[3] = j.run_len;
    z.n[3] = (uint64_t)set.groups.size() * j.count;
    z.count = 4;
    HIP_TRY(bpx::launch_zero_u32(z, st));
    unsigned long long* ustats = j.meta + kMetaWords;
    uint32_t* overflow = reinterpret_cast<uint32_t*>(j.meta + 1);
    if (j.ev_start) HIP_TRY(hipEventRecord(j.ev_start, st));
    uint64_t seg_at = 0;
    const int groups = j.ctx->num_cus * j.groups_per_cu;
    for (size_t ci = 0; ci < j.nclasses; ++ci) {
        const auto& c = j.classes[ci];
        bpx::ScanParams sp{};
        sp.residues = j.db->d_arena;
        sp.soff = j.db->d_off;
        sp.slen = j.db->d_len;
        sp.order = j.order;
        sp.poff = j.poff;
        sp.plen = j.plen;
        sp.first = j.first;
        sp.count = j.count;
        sp.groups = set.d_groups;
        sp.qinfo = set.d_qinfo;
        sp.tables = set.d_tables;
        sp.qres = set.d_qres;
        sp.qmap = set.d_qmap;
        sp.matrix = set.d_matrix;
        sp.ngroups_pv = set.ngroups_pv;
        sp.word = set.word;
        sp.window = set.window;
        sp.items = j.items + c.item_first;
        sp.nitems = c.nitems;
        sp.flat = c.flat;
        sp.flat_counters = j.work + kMaxClasses + c.item_first;
        sp.work_counter = j.work + ci;
        sp.cnt = j.cnt;
        sp.stats = j.sstats;
        sp.run_loc = j.run_loc;
        sp.run_len = j.run_len;
        sp.ustats = ustats;
        // each launch class owns its own range of wave segments; run
        // locations written by the kernel are relative to sp.seg, so they
        // carry the class offset below
        sp.seg = j.seg;
        sp.seg_cap = j.seg_cap;
        sp.seg_wave_base = (uint32_t)(seg_at / j.seg_cap);
        sp.overflow = overflow;
        sp.phase = j.phase;
        sp.max_fill = reinterpret_cast<uint32_t*>(j.meta + 2);
        sp.lds_table = c.plan.table;
        sp.lds_qres = c.plan.qres;
        sp.lds_qmap = c.plan.qmap;
        sp.lds_qinfo = c.plan.qinfo;
        sp.lds_matrix = c.plan.matrix;
        sp.lds_wave = c.plan.wave;
        sp.wave_stride = c.plan.wave_stride;
        sp.lds_total = c.plan.total;
        sp.table_cap_words = c.plan.table_cap_words;
        sp.qres_cap = c.plan.qres_cap;
        sp.max_nq = c.plan.max_nq;
        sp.diag_cap = c.diag_cap;
        sp.qcap = c.plan.qcap;
        sp.claim_slots = c.plan.claim_slots;
        sp.sbuf = j.sbuf;
        HIP_TRY(bpx::launch_scan(sp, c.tab_lds, c.multi, c.narrow, c.waves, groups, st, !j.capturing));
        seg_at += (uint64_t)groups * c.waves * j.seg_cap;
    }
    if (j.ev_kernels) HIP_TRY(hipEventRecord(j.ev_kernels, st));
    bpx::CompactParams cp{};
    cp.cnt = j.cnt;
    cp.n = (uint64_t)set.nq * j.count;
    cp.tile_sum = j.tile;
    cp.run_loc = j.run_loc;
    cp.run_len = j.run_len;
    cp.groups = set.d_groups;
    cp.qinfo = set.d_qinfo;
    cp.ngroups = (uint32_t)set.groups.size();
    cp.count = j.count;
    cp.first = j.first;
    cp.seg = j.seg;
    cp.records = j.records;
    cp.subjects = j.subjects;
    cp.queries = j.queries;
    cp.rec_cap = j.rec_cap;
    cp.total = j.meta;
    cp.overflow = overflow;
    HIP_TRY(bpx::launch_compact(cp, st));
    HIP_TRY(hipMemcpyAsync(j.h_meta, j.meta, (kMetaWords + 5ull * set.nq) * 8, hipMemcpyDeviceToHost, st));
    return BPX_OK;
}

// The job of a pooled result.  `events` brackets the word-finder launches
// with the result's timing events (not when captured into a graph: then the
// caller brackets the graph launch).
ScanJob result_job(bpx_result* r, const uint32_t* order, bool events, bool start_event) {
    ScanJob j;
    j.ctx = r->ctx;
    j.set = r->set;
    j.db = r->db;
    j.first = r->first;
    j.count = r->count;
    j.classes = r->classes.data();
    j.nclasses = r->classes.size();
    j.nitems = r->items.size();
    j.groups_per_cu = r->launch.groups_per_cu;
    j.sbuf = (uint32_t)r->launch.subject_buffer;
    j.stream = r->stream;
    j.order = order;
    j.poff = r->d_poff ? r->d_poff : r->db->d_poff;
    j.plen = r->d_plen ? r->d_plen : r->db->d_plen;
    j.cnt = r->d_cnt;
    j.sstats = r->d_sstats;
    j.run_loc = r->d_run_loc;
    j.run_len = r->d_run_len;
    j.seg = r->d_seg;
    j.seg_cap = r->seg_cap;
    j.tile = r->d_tile;
    j.records = r->d_records;
    j.subjects = r->d_subjects;
    j.queries = r->d_queries;
    j.meta = r->d_meta;
    j.items = r->d_items;
    j.work = r->d_work;
    j.h_meta = r->h_meta;
    j.phase = r->d_phase;
    if (events) {
        if (start_event) j.ev_start = r->ev[0];
        j.ev_kernels = r->ev[1];
    }
    j.capturing = !events;
    return j;
}

int upload_items(bpx_result* r) {
    if (!r->items.empty())
        HIP_TRY(hipMemcpyAsync(r->d_items, r->items.data(), r->items.size() * sizeof(bpx::WorkItem),
                               hipMemcpyHostToDevice, r->stream));
    return BPX_OK;
}

// start_event = false keeps the scan's first start event: a rescan's time
// then includes the launch whose segments overflowed
int enqueue(bpx_result* r, const uint32_t* order, bool start_event = true) {
    int rc = upload_items(r);
    return rc ? rc : enqueue_job(result_job(r, order, true, start_event));
}

void graph_release(bpx_result* r) {
    if (r->graph_exec) (void)hipGraphExecDestroy(r->graph_exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    r->graph_exec = nullptr;
    r->graph = nullptr;
}

// Captures the device work of the scan into a HIP graph and launches it,
// bracketed by the timing events.
int graph_capture_and_launch(bpx_result* r) {
    const uint32_t* order = r->d_order ? r->d_order : r->db->d_order;
    graph_release(r);
    int rc = upload_items(r);
    if (rc) return rc;
    // everything that allocates or sets kernel attributes happens before the
    // capture (the buffers, in alloc_buffers, and each launch's LDS size)
    for (const auto& c : r->classes) {
        bpx::ScanParams sp{};
        sp.word = r->set->word;
        sp.lds_total = c.plan.total;
        HIP_TRY(bpx::prepare_scan(sp, c.tab_lds, c.multi, c.narrow));
    }
    HIP_TRY(hipStreamSynchronize(r->stream));
    HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_job(result_job(r, order, false, false));
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(r->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(BPX_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
    r->graph = g;
    HIP_TRY(hipGraphInstantiate(&r->graph_exec, r->graph, nullptr, nullptr, 0));
    r->graph_stale = false;
    HIP_TRY(hipEventRecord(r->ev[0], r->stream));
    HIP_TRY(hipGraphLaunch(r->graph_exec, r->stream));
    HIP_TRY(hipEventRecord(r->ev[1], r->stream));
    return BPX_OK;
}

// Subjects of [first, first + count) too long for narrow diagonal slots.
// They lead the longest-first processing order.
uint32_t count_wide(const bpx_db* db, uint32_t first, uint32_t count, int32_t window) {
    uint32_t n = 0;
    for (uint32_t id : db->h_order) {
        if ((uint64_t)db->h_len[id] + (uint64_t)std::max(window, 0) < bpx::kNarrowLimit) break;
        if (id >= first && id - first < count) ++n;
    }
    return n;
}

int start_scan(bpx_ctx* ctx, const QuerySet& set, const bpx_db* db, uint32_t first, uint32_t count,
               const bpx_launch* launch, uint32_t out_flags, bpx_result** out) {
    if (!ctx || !db || !out) return fail(BPX_E_ARG, "NULL handle");
    *out = nullptr;
    if (set.ctx != ctx || db->ctx != ctx) return fail(BPX_E_ARG, "handles belong to another context");
    if ((uint64_t)first + count > db->n) return fail(BPX_E_ARG, "subject range outside the database");
    HIP_TRY(hipSetDevice(ctx->device));
    auto* r = new (std::nothrow) bpx_result;
    if (!r) return fail(BPX_E_NOMEM, "host allocation failed");
    r->ctx = ctx;
    r->set = &set;
    r->db = db;
    r->first = first;
    r->count = count;
    r->flags = out_flags;
    r->stream = ctx->lanes[ctx->next_lane++ & 1u];
    if (db->cost_order && set.window > kCostMaxWindow) {
        delete r;
        return fail(BPX_E_ARG, "window %d above %d while the database has a cost order", set.window, kCostMaxWindow);
    }
    int rc = plan_scan(r, set, count, count_wide(db, first, count, set.window), launch);
    if (rc) {
        delete r;
        return rc;
    }
    for (auto& ev : r->ev)
        if (hipEventCreate(&ev) != hipSuccess) {
            bpx_result_destroy(r);
            return fail(BPX_E_HIP, "event creation failed");
        }
    const uint64_t waves = std::max<uint64_t>(total_waves(r), 1);
    const uint64_t pairs = (uint64_t)set.nq * count;
    const uint32_t seg_cap =
        (uint32_t)std::max<uint64_t>(std::max<uint64_t>(kSegFloor, ctx->seg_hint), (pairs / 8 + 65536) / waves + 1);
    rc = alloc_buffers(r, seg_cap);
    if (rc) {
        bpx_result_destroy(r);
        return rc;
    }
    const uint32_t* order = db->d_order;
    if (!(first == 0 && count == db->n)) {
        std::vector<uint32_t> sub;
        sub.reserve(count);
        for (uint32_t id : db->h_order)
            if (id >= first && id - first < count) sub.push_back(id);
        std::vector<uint64_t> soff(sub.size());
        std::vector<uint32_t> slen(sub.size());
        for (size_t k = 0; k < sub.size(); ++k) {
            soff[k] = db->h_off[sub[k]];
            slen[k] = db->h_len[sub[k]];
        }
        r->d_order = (uint32_t*)ctx->pool.get(std::max<size_t>(count, 1) * 4);
        r->d_poff = (uint64_t*)ctx->pool.get(std::max<size_t>(count, 1) * 8);
        r->d_plen = (uint32_t*)ctx->pool.get(std::max<size_t>(count, 1) * 4);
        if (!r->d_order || !r->d_poff || !r->d_plen) {
            bpx_result_destroy(r);
            return fail(BPX_E_NOMEM, "work list allocation failed");
        }
        if (count) {
            hipError_t e = hipMemcpyAsync(r->d_order, sub.data(), (size_t)count * 4, hipMemcpyHostToDevice, r->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(r->d_poff, soff.data(), (size_t)count * 8, hipMemcpyHostToDevice, r->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(r->d_plen, slen.data(), (size_t)count * 4, hipMemcpyHostToDevice, r->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
            if (e != hipSuccess) {
                bpx_result_destroy(r);
                return fail(BPX_E_HIP, "work list upload failed: %s", hipGetErrorString(e));
            }
        }
        order = r->d_order;
    }
    rc = (out_flags & BPX_OUT_GRAPH) ? graph_capture_and_launch(r) : enqueue(r, order);
    if (rc) {
        bpx_result_destroy(r);
        return rc;
    }
    *out = r;
    return BPX_OK;
}

}  // namespace

namespace {
// the processing order and the subject descriptors in that order
hipError_t upload_order(bpx_db* db) {
    const size_t n = db->n;
    if (n == 0) return hipSuccess;
    std::vector<uint64_t> poff(n);
    std::vector<uint32_t> plen(n);
    for (size_t k = 0; k < n; ++k) {
        poff[k] = db->h_off[db->h_order[k]];
        plen[k] = db->h_len[db->h_order[k]];
    }
    hipError_t e = hipMemcpy(db->d_order, db->h_order.data(), n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db->d_poff, poff.data(), n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db->d_plen, plen.data(), n * 4, hipMemcpyHostToDevice);
    return e;
}

// Subject ids longest first, ties by id: the dynamic work queue then behaves
// like longest-processing-time-first list scheduling.  A counting sort by
// length (stable), or a stable comparison sort for very long subjects.
std::vector<uint32_t> longest_first(const std::vector<uint32_t>& len, uint32_t maxlen) {
    const size_t n = len.size();
    std::vector<uint32_t> order(n);
    if (maxlen <= (1u << 22)) {
        std::vector<uint32_t> start((size_t)maxlen + 2, 0);
        for (uint32_t l : len) ++start[(size_t)(maxlen - l) + 1];
        for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
        for (size_t i = 0; i < n; ++i) order[start[maxlen - len[i]]++] = (uint32_t)i;
    } else {
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return len[a] > len[b]; });
    }
    return order;
}

// A host array to copy to the device through the staging buffers.
struct HostCopy {
    void* dst;
    const void* src;
    uint64_t bytes;
};

// Repacks the residues into the arena layout and uploads it, then the
// descriptor arrays: chunks of whole subjects are filled by host threads into
// one of two pinned staging buffers while the other one's copy runs.
// Residue codes are checked on the way; the first bad one (lowest subject,
// then position) is reported.
int upload_arena(bpx_ctx* ctx, uint8_t* d_arena, const uint8_t* residues, const uint64_t* offsets,
                 const uint32_t* lengths, const std::vector<uint64_t>& off, uint64_t arena,
                 const std::vector<HostCopy>& extra) {
    const uint32_t n = (uint32_t)off.size();
    const uint64_t kChunk = 64ull << 20;
    uint64_t biggest = kArenaTail, need = arena;
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t end = i + 1 < n ? off[i + 1] : arena;
        biggest = std::max(biggest, end - off[i]);
    }
    for (const auto& h : extra) need = std::max(need, h.bytes);
    const uint64_t cap = std::max(biggest, std::min(kChunk, need));
    // staging buffers of its own (not the context's pools, which a server
    // thread on the same context may be using)
    uint8_t* stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    int rc = BPX_OK;
    for (auto& p : stage)
        if (rc == BPX_OK && hipHostMalloc((void**)&p, cap, hipHostMallocDefault) != hipSuccess)
            rc = fail(BPX_E_NOMEM, "pinned staging allocation of %llu bytes failed", (unsigned long long)cap);
    for (auto& ev : done)
        if (rc == BPX_OK && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            rc = fail(BPX_E_HIP, "event creation failed");
    uint64_t c = 0;  // copies submitted; copy c uses stage[c & 1]
    auto next_buf = [&]() -> uint8_t* {
        if (c >= 2 && hipEventSynchronize(done[c & 1]) != hipSuccess) return nullptr;
        return stage[c & 1];
    };
    auto submit = [&](void* dst, uint64_t bytes) -> bool {
        if (hipMemcpyAsync(dst, stage[c & 1], bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
            hipEventRecord(done[c & 1], ctx->stream) != hipSuccess)
            return false;
        ++c;
        return true;
    };
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    uint32_t i0 = 0;
    while (rc == BPX_OK) {
        // whole subjects [i0, i1) of at most cap bytes; the last chunk ends at the arena's end
        uint32_t i1 = i0;
        const uint64_t base = i0 < n ? off[i0] : 0;
        while (i1 < n && (i1 + 1 < n ? off[i1 + 1] : arena) - base <= cap) ++i1;
        if (i1 == i0 && i0 < n) i1 = i0 + 1;
        const uint64_t top = i1 < n ? off[i1] : arena;
        uint8_t* buf = next_buf();
        if (!buf) {
            rc = fail(BPX_E_HIP, "arena upload failed");
            break;
        }
        std::vector<uint64_t> bad(nthreads, ~0ull);  // (subject << 32 | position) of each thread's first bad code
        auto work = [&](unsigned t) {
            const uint32_t a = i0 + (uint32_t)((uint64_t)(i1 - i0) * t / nthreads);
            const uint32_t b = i0 + (uint32_t)((uint64_t)(i1 - i0) * (t + 1) / nthreads);
            for (uint32_t i = a; i < b; ++i) {
                const uint8_t* __restrict__ src = residues + offsets[i];
                uint8_t* __restrict__ dst = buf + (off[i] - base);
                const uint32_t L = lengths[i];
                unsigned flag = 0;
                for (uint32_t k = 0; k < L; ++k) {  // copy and check in one pass
                    const uint8_t v = src[k];
                    dst[k] = v;
                    flag |= v >= BPX_ALPHABET;
                }
                if (flag && bad[t] == ~0ull) {
                    uint32_t k = 0;
                    while (src[k] < BPX_ALPHABET) ++k;
                    bad[t] = (uint64_t)i << 32 | k;
                }
                const uint64_t end = (i + 1 < n ? off[i + 1] : arena) - off[i];
                std::memset(dst + L, 0, end - L);
            }
        };
        if (i0 == n) {
            std::memset(buf, 0, arena - base);  // no subjects: the zero tail alone
        } else if (i1 - i0 < 64 * nthreads) {
            for (unsigned t = 0; t < nthreads; ++t) work(t);
        } else {
            // partitions a thread could not be started for run here
            std::vector<std::thread> pool;
            std::vector<unsigned> here{0};
            for (unsigned t = 1; t < nthreads; ++t) {
                try {
                    pool.emplace_back(work, t);
                } catch (...) {
                    here.push_back(t);
                }
            }
            for (unsigned t : here) work(t);
            for (auto& th : pool) th.join();
        }
        const uint64_t first_bad = *std::min_element(bad.begin(), bad.end());
        if (first_bad != ~0ull) {
            const uint32_t i = (uint32_t)(first_bad >> 32), k = (uint32_t)first_bad;
            rc = fail(BPX_E_ARG, "subject %u position %u: residue code %u >= 28", i, k, residues[offsets[i] + k]);
            break;
        }
        if (!submit(d_arena + base, top - base)) {
            rc = fail(BPX_E_HIP, "arena upload failed");
            break;
        }
        i0 = i1;
        if (i0 >= n) break;
    }
    for (const auto& h : extra)
        for (uint64_t at = 0; rc == BPX_OK && at < h.bytes; at += cap) {
            const uint64_t b = std::min(cap, h.bytes - at);
            uint8_t* buf = next_buf();
            if (buf) std::memcpy(buf, static_cast<const uint8_t*>(h.src) + at, b);
            if (!buf || !submit(static_cast<uint8_t*>(h.dst) + at, b)) rc = fail(BPX_E_HIP, "descriptor upload failed");
        }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && rc == BPX_OK) rc = fail(BPX_E_HIP, "arena upload failed");
    for (auto& ev : done)
        if (ev) (void)hipEventDestroy(ev);
    for (auto* p : stage)
        if (p) (void)hipHostFree(p);
    return rc;
}
}  // namespace

// ---- the allocation-free scan's scratch ------------------------------------------
//
// One file-scope device array, a copy per GPU, holds every buffer of
// bpx_scan_noalloc beyond its inputs and outputs, at fixed offsets.  Host
// code never uses the array's name as an address: each context keeps its
// device's address from hipGetSymbolAddress.
namespace {
constexpr uint64_t kNaMeta = 0;                                         // kMetaWords + 5 u64
constexpr uint64_t kNaWork = 256;                                       // kMaxClasses + 2 u32
constexpr uint64_t kNaItems = kNaWork + 1024;                           // 2 WorkItem
constexpr uint64_t kNaTile = kNaItems + 256;                            // tile sums
constexpr uint64_t kNaCnt = kNaTile + BPX_NOALLOC_MAX_SUBJECTS / bpx::kTile * 4ull;
constexpr uint64_t kNaRunLoc = kNaCnt + BPX_NOALLOC_MAX_SUBJECTS * 4ull;
constexpr uint64_t kNaRunLen = kNaRunLoc + BPX_NOALLOC_MAX_SUBJECTS * 4ull;
constexpr uint64_t kNaSeg = kNaRunLen + BPX_NOALLOC_MAX_SUBJECTS * 4ull;
constexpr uint64_t kNaBytes = kNaSeg + BPX_NOALLOC_SEG_RECORDS * (uint64_t)sizeof(bpx::SegRecord);
static_assert((kMetaWords + 5) * 8 <= kNaWork && (kMaxClasses + 2) * 4 <= kNaItems - kNaWork &&
                  2 * sizeof(bpx::WorkItem) <= kNaTile - kNaItems,
              "scratch regions overlap");
static_assert(kNaCnt % 256 == 0 && kNaSeg % 256 == 0, "scratch regions are 256-byte aligned");
static_assert(kNaBytes == BPX_NOALLOC_SCRATCH_BYTES, "BPX_NOALLOC_SCRATCH_BYTES is the scratch size");
}  // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};

// The scratch array's address on `device`, which must be current: resolved
// once per device, then reused by every context on it.
hipError_t scratch_address(int device, uint8_t** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[device] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}
}  // namespace

namespace bpx {
void set_last_error(const char* msg) { std::snprintf(g_err, sizeof g_err, "%s", msg ? msg : ""); }
}  // namespace bpx

extern "C" {

const char* bpx_last_error(void) { return g_err; }

const int8_t* bpx_blosum62(void) { return bpx::blosum62(); }

int bpx_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

int bpx_context_create(int device, bpx_ctx** out) {
    if (!out) return fail(BPX_E_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(BPX_E_NODEV, "no HIP device");
    if (device < 0 || device >= n) return fail(BPX_E_ARG, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    auto* ctx = new (std::nothrow) bpx_ctx;
    if (!ctx) return fail(BPX_E_NOMEM, "host allocation failed");
    ctx->device = device;
    int v = 0;
    hipError_t e = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device);
    ctx->num_cus = (e == hipSuccess && v > 0) ? v : 1;
    e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    ctx->lds_limit = (e == hipSuccess && v > 0) ? v : 65536;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) ctx->hbm_bytes = (int64_t)tot;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->lanes[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->lanes[1], hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(BPX_E_HIP, "stream / counter allocation failed");
    }
    // the allocation-free scan's scratch on this device (current); a failure
    // is reported by bpx_scan_noalloc, the only user
    e = scratch_address(device, &ctx->scratch);
    if (e != hipSuccess) {
        std::snprintf(ctx->scratch_err, sizeof ctx->scratch_err, "hipGetSymbolAddress(scratch) on device %d: %s",
                      device, hipGetErrorString(e));
        (void)hipGetLastError();  // kept for bpx_scan_noalloc, not for the next launch check
    }
    // BPX_POOL_FAIL_FIRST=n (tests): the first n device-pool allocations of
    // this context fail once before their retry
    if (const char* f = std::getenv("BPX_POOL_FAIL_FIRST")) ctx->pool.fail_first = std::max(0, std::atoi(f));
    // the attribute and memory queries above fall back to defaults on
    // failure: none of their errors may reach a later launch check
    (void)hipGetLastError();
    *out = ctx;
    return BPX_OK;
}

int bpx_context_destroy(bpx_ctx* ctx) {
    if (!ctx) return BPX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t l : ctx->lanes)
        if (l) {
            (void)hipStreamSynchronize(l);
            (void)hipStreamDestroy(l);
        }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BPX_OK;
}

int bpx_context_info(const bpx_ctx* ctx, int32_t* num_cus, int64_t* lds_bytes, int64_t* hbm_bytes) {
    if (!ctx) return fail(BPX_E_ARG, "ctx is NULL");
    if (num_cus) *num_cus = ctx->num_cus;
    if (lds_bytes) *lds_bytes = ctx->lds_limit;
    if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
    return BPX_OK;
}

int bpx_db_load(bpx_ctx* ctx, const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                uint32_t num_subjects, bpx_db** out) {
    if (!ctx || !out) return fail(BPX_E_ARG, "ctx/out is NULL");
    *out = nullptr;
    if (num_subjects && (!offsets || !lengths || !residues)) return fail(BPX_E_ARG, "NULL arrays");
    HIP_TRY(hipSetDevice(ctx->device));
    // The device layout: 128-byte-aligned subject starts, zero padding, and
    // a readable zero tail.
    std::vector<uint64_t> off(num_subjects);
    uint64_t at = 0, nres = 0;
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < num_subjects; ++i) {
        if (lengths[i] >= (1u << 28)) return fail(BPX_E_ARG, "subject %u longer than 2^28 residues", i);
        off[i] = at;
        at += ((uint64_t)lengths[i] + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        nres += lengths[i];
        maxlen = std::max(maxlen, lengths[i]);
    }
    const uint64_t arena = at + kArenaTail;
    auto* db = new (std::nothrow) bpx_db;
    if (!db) return fail(BPX_E_NOMEM, "host allocation failed");
    db->ctx = ctx;
    db->n = num_subjects;
    db->nres = nres;
    db->arena_bytes = arena;
    db->h_len.assign(lengths, lengths + num_subjects);
    db->h_order = longest_first(db->h_len, maxlen);
    const size_t n1 = std::max<uint32_t>(num_subjects, 1);
    if (hipMalloc(&db->d_arena, arena) != hipSuccess || hipMalloc(&db->d_off, n1 * 8) != hipSuccess ||
        hipMalloc(&db->d_len, n1 * 4) != hipSuccess || hipMalloc(&db->d_order, n1 * 4) != hipSuccess ||
        hipMalloc(&db->d_poff, n1 * 8) != hipSuccess || hipMalloc(&db->d_plen, n1 * 4) != hipSuccess) {
        bpx_db_destroy(db);
        return fail(BPX_E_NOMEM, "device allocation of %llu-byte arena failed", (unsigned long long)arena);
    }
    // descriptors by subject id and, in processing order, the order itself
    // with each subject's offset and length
    std::vector<uint64_t> poff(num_subjects);
    std::vector<uint32_t> plen(num_subjects);
    for (uint32_t k = 0; k < num_subjects; ++k) {
        poff[k] = off[db->h_order[k]];
        plen[k] = lengths[db->h_order[k]];
    }
    const std::vector<HostCopy> extra = {{db->d_off, off.data(), (uint64_t)num_subjects * 8},
                                         {db->d_len, lengths, (uint64_t)num_subjects * 4},
                                         {db->d_order, db->h_order.data(), (uint64_t)num_subjects * 4},
                                         {db->d_poff, poff.data(), (uint64_t)num_subjects * 8},
                                         {db->d_plen, plen.data(), (uint64_t)num_subjects * 4}};
    int rc = upload_arena(ctx, db->d_arena, residues, offsets, lengths, off, arena, extra);
    if (rc) {
        bpx_db_destroy(db);
        return rc;
    }
    db->h_off = std::move(off);
    *out = db;
    return BPX_OK;
}

int bpx_db_set_cost(bpx_db* db, const float* cost) {
    if (!db) return fail(BPX_E_ARG, "db is NULL");
    HIP_TRY(hipSetDevice(db->ctx->device));
    std::vector<uint32_t> order(db->n);
    std::iota(order.begin(), order.end(), 0u);
    if (cost) {
        for (uint32_t i = 0; i < db->n; ++i)
            if (!(cost[i] >= 0.f) || !std::isfinite(cost[i]))
                return fail(BPX_E_ARG, "cost[%u] is negative or not finite", i);
        // long subjects (wide diagonal slots) first, longest first; then by cost
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const bool la = db->h_len[a] >= kCostLongPrefix, lb = db->h_len[b] >= kCostLongPrefix;
            if (la != lb) return la;
            if (la) return db->h_len[a] > db->h_len[b];
            return cost[a] > cost[b];
        });
    } else {
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return db->h_len[a] > db->h_len[b]; });
    }
    // Bins of equal predicted work along the order: a scan of the whole
    // database gives each its own work item, whose wavefronts claim subjects
    // in chunks sized by the bin's mean cost (the heaviest one at a time),
    // so a claim holds about the same work in every bin.
    std::vector<uint32_t> bin_end;
    std::vector<float> bin_weight;
    if (cost && db->n) {
        double total = 0;
        for (uint32_t i = 0; i < db->n; ++i) total += cost[i];
        if (total > 0) {
            const double mean = total / db->n;
            double acc = 0, bin_sum = 0;
            uint32_t bin_first = 0;
            for (uint32_t k = 0; k < db->n; ++k) {
                const double c = cost[order[k]];
                acc += c;
                bin_sum += c;
                const bool cut = acc >= total * (double)(bin_end.size() + 1) / kCostBins || k + 1 == db->n;
                if (cut) {
                    bin_end.push_back(k + 1);
                    bin_weight.push_back((float)(bin_sum / (k + 1 - bin_first) / mean));
                    bin_first = k + 1;
                    bin_sum = 0;
                }
            }
        }
    }
    // Within a bin, subjects go longest first: its subjects cost about the
    // same, and neighbours of similar length share a wavefront's subject
    // buffer (a pack) as in the length order.
    for (size_t b = 0; b < bin_end.size(); ++b) {
        const uint32_t f = b ? bin_end[b - 1] : 0u;
        std::stable_sort(order.begin() + f, order.begin() + bin_end[b], [&](uint32_t x, uint32_t y) {
            const bool lx = db->h_len[x] >= kCostLongPrefix, ly = db->h_len[y] >= kCostLongPrefix;
            if (lx != ly) return lx;
            return db->h_len[x] > db->h_len[y];
        });
    }
    db->h_order.swap(order);
    HIP_TRY(upload_order(db));
    db->cost_order = cost != nullptr;
    db->bin_end.swap(bin_end);
    db->bin_weight.swap(bin_weight);
    return BPX_OK;
}

int bpx_db_features(bpx_ctx* ctx, const bpx_db* db, const uint16_t* word_counts, int32_t word_length,
                    float* features) {
    if (!ctx || !db || !word_counts || !features) return fail(BPX_E_ARG, "NULL argument");
    if (word_length != 2 && word_length != 3) return fail(BPX_E_ARG, "word_length must be 2 or 3");
    if (db->ctx != ctx) return fail(BPX_E_ARG, "database belongs to another context");
    if (db->n == 0) return BPX_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nw = (size_t)1 << (5 * word_length);
    uint16_t* d_counts = (uint16_t*)ctx->pool.get(nw * 2);
    float* d_out = (float*)ctx->pool.get((size_t)db->n * BPX_CM_FEATURES * 4);
    if (!d_counts || !d_out) {
        ctx->pool.put(d_counts);
        ctx->pool.put(d_out);
        return fail(BPX_E_NOMEM, "feature buffers");
    }
    hipError_t e = hipMemcpyAsync(d_counts, word_counts, nw * 2, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = bpx::launch_features(db->d_arena, db->d_off, db->d_len, db->n, d_counts, word_length, d_out, ctx->num_cus,
                                 ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(features, d_out, (size_t)db->n * BPX_CM_FEATURES * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->pool.put(d_counts);
    ctx->pool.put(d_out);
    if (e != hipSuccess) return fail(BPX_E_HIP, "features: %s", hipGetErrorString(e));
    return BPX_OK;
}

int bpx_db_info(const bpx_db* db, uint32_t* num_subjects, uint64_t* num_residues, uint64_t* arena_bytes) {
    if (!db) return fail(BPX_E_ARG, "db is NULL");
    if (num_subjects) *num_subjects = db->n;
    if (num_residues) *num_residues = db->nres;
    if (arena_bytes) *arena_bytes = db->arena_bytes;
    return BPX_OK;
}

int bpx_db_destroy(bpx_db* db) {
    if (!db) return BPX_OK;
    (void)hipSetDevice(db->ctx->device);
    (void)hipFree(db->d_arena);
    (void)hipFree(db->d_off);
    (void)hipFree(db->d_len);
    (void)hipFree(db->d_order);
    (void)hipFree(db->d_poff);
    (void)hipFree(db->d_plen);
    delete db;
    return BPX_OK;
}

int bpx_query_prepare(bpx_ctx* ctx, const uint8_t* query, int32_t qlen, const int8_t* matrix,
                      const bpx_word_params* params, bpx_query** out) {
    if (!ctx || !out || !params) return fail(BPX_E_ARG, "ctx/params/out is NULL");
    *out = nullptr;
    if (qlen < 0 || qlen > 65535 || (qlen > 0 && !query))
        return fail(BPX_E_ARG, "query length %d outside [0, 65535]", qlen);
    bpx_word_params p = *params;
    if (p.word_length != 2 && p.word_length != 3) return fail(BPX_E_ARG, "word_length must be 2 or 3");
    if (p.window <= 0 || p.x_dropoff <= 0) return fail(BPX_E_ARG, "window and x_dropoff must be positive");
    if (p.diag_mask == 0) p.diag_mask = next_pow2((int64_t)qlen + p.window) - 1;
    if (p.diag_mask < 0 || ((p.diag_mask + 1) & p.diag_mask) != 0 || p.diag_mask + 1 > (1 << 17))
        return fail(BPX_E_ARG, "diag_mask must be 2^k - 1 with 2^k <= 131072");
    const int8_t* M = matrix ? matrix : bpx::blosum62();
    auto* q = new (std::nothrow) bpx_query;
    if (!q) return fail(BPX_E_NOMEM, "host allocation failed");
    q->qlen = qlen;
    q->params = p;
    std::vector<QuerySpec> qs{{query, qlen, p.x_dropoff, p.cutoff_score, p.diag_mask + 1}};
    std::vector<bpx::HostTable> tabs;
    int rc = build_set(ctx, qs, M, p.word_length, p.threshold, p.window, 0, 1, 0, &q->set, &tabs);
    if (rc) {
        delete q;
        return rc;
    }
    q->npresent = (int32_t)tabs[0].npresent;
    q->nentries = (int64_t)tabs[0].entries.size();
    *out = q;
    return BPX_OK;
}

int bpx_query_info(const bpx_query* q, int64_t* table_bytes, int32_t* lds_resident, int32_t* num_present_words,
                   int64_t* num_entries) {
    if (!q) return fail(BPX_E_ARG, "query is NULL");
    if (table_bytes) *table_bytes = (int64_t)q->set.table_bytes;
    if (lds_resident) {
        bpx_result tmp;
        tmp.ctx = q->set.ctx;
        *lds_resident = 0;
        if (plan_scan(&tmp, q->set, 1, 0, nullptr) == BPX_OK && !tmp.classes.empty())
            *lds_resident = tmp.classes[0].tab_lds ? 1 : 0;
    }
    if (num_present_words) *num_present_words = q->npresent;
    if (num_entries) *num_entries = q->nentries;
    return BPX_OK;
}

int bpx_query_destroy(bpx_query* q) {
    if (!q) return BPX_OK;
    q->set.release();
    delete q;
    return BPX_OK;
}

int bpx_batch_prepare(bpx_ctx* ctx, uint32_t num_queries, const uint8_t* const* queries, const int32_t* qlens,
                      const int8_t* matrix, const bpx_word_params* params, const int32_t* x_dropoff,
                      const int32_t* cutoff, const bpx_batch_policy* policy, bpx_batch** out) {
    if (!ctx || !out || !params) return fail(BPX_E_ARG, "ctx/params/out is NULL");
    *out = nullptr;
    if (num_queries && (!queries || !qlens)) return fail(BPX_E_ARG, "NULL query arrays");
    const bpx_word_params& p = *params;
    if (p.word_length != 2 && p.word_length != 3) return fail(BPX_E_ARG, "word_length must be 2 or 3");
    if (p.window <= 0) return fail(BPX_E_ARG, "window must be positive");
    if (p.diag_mask != 0) return fail(BPX_E_ARG, "a batch derives diag_mask per query; pass 0");
    std::vector<QuerySpec> qs(num_queries);
    for (uint32_t i = 0; i < num_queries; ++i) {
        if (qlens[i] < 0 || qlens[i] > 65535 || (qlens[i] > 0 && !queries[i]))
            return fail(BPX_E_ARG, "query %u: length %d outside [0, 65535]", i, qlens[i]);
        qs[i].seq = queries[i];
        qs[i].len = qlens[i];
        qs[i].xdrop = x_dropoff ? x_dropoff[i] : p.x_dropoff;
        qs[i].cutoff = cutoff ? cutoff[i] : p.cutoff_score;
        if (qs[i].xdrop <= 0) return fail(BPX_E_ARG, "query %u: x_dropoff must be positive", i);
        qs[i].diag_len = next_pow2((int64_t)qlens[i] + p.window);
    }
    int32_t gres = policy && policy->group_residues > 0 ? policy->group_residues : 600;
    const int32_t gq = policy && policy->group_queries > 0 ? std::min(policy->group_queries, 255) : 32;
    gres = std::min(gres, 65535);
    auto* b = new (std::nothrow) bpx_batch;
    if (!b) return fail(BPX_E_NOMEM, "host allocation failed");
    const int32_t gd = policy && policy->group_diag_slots > 0 ? policy->group_diag_slots : 0;
    int rc = build_set(ctx, qs, matrix ? matrix : bpx::blosum62(), p.word_length, p.threshold, p.window, gres, gq, gd,
                       &b->set);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return BPX_OK;
}

int bpx_batch_info(const bpx_batch* b, uint32_t* num_queries, uint32_t* num_groups, int64_t* table_bytes) {
    if (!b) return fail(BPX_E_ARG, "batch is NULL");
    if (num_queries) *num_queries = b->set.nq;
    if (num_groups) *num_groups = (uint32_t)b->set.groups.size();
    if (table_bytes) *table_bytes = (int64_t)b->set.table_bytes;
    return BPX_OK;
}

int bpx_batch_destroy(bpx_batch* b) {
    if (!b) return BPX_OK;
    b->set.release();
    delete b;
    return BPX_OK;
}

int bpx_check_memory(bpx_ctx* ctx, const bpx_db* db, const bpx_query* q, uint32_t first, uint32_t count,
                     const bpx_launch* launch, size_t* device_bytes, size_t* lds_bytes) {
    if (!ctx || !db || !q) return fail(BPX_E_ARG, "NULL handle");
    if ((uint64_t)first + count > db->n) return fail(BPX_E_ARG, "subject range outside the database");
    bpx_result tmp;
    tmp.ctx = ctx;
    int rc = plan_scan(&tmp, q->set, count, count_wide(db, first, count, q->set.window), launch);
    const uint32_t lds = tmp.lds_need;
    const uint64_t waves = std::max<uint64_t>(total_waves(&tmp), 1);
    const uint64_t seg = std::max<uint64_t>(kSegFloor, ((uint64_t)count * q->set.nq / 8 + 65536) / waves + 1) * waves;
    const uint64_t need = db->arena_bytes + (uint64_t)db->n * 16 + q->set.table_bytes + (uint64_t)count * 24 +
                          seg * (sizeof(bpx::SegRecord) + sizeof(bpx_init_hsp) + 8);
    if (device_bytes) *device_bytes = (size_t)need;
    if (lds_bytes) *lds_bytes = lds;
    if (rc) return rc;
    if (ctx->hbm_bytes && need > (uint64_t)ctx->hbm_bytes)
        return fail(BPX_E_BUDGET, "needs %llu device bytes", (unsigned long long)need);
    return BPX_OK;
}

int bpx_context_trim(bpx_ctx* ctx) {
    if (!ctx) return fail(BPX_E_ARG, "ctx is NULL");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (hipStream_t l : ctx->lanes) HIP_TRY(hipStreamSynchronize(l));
    ctx->pool.trim();
    ctx->pinned.trim();
    ctx->coherent.trim();
    return BPX_OK;
}

int bpx_context_limit(bpx_ctx* ctx, int64_t hbm_bytes) {
    if (!ctx || hbm_bytes < 0) return fail(BPX_E_ARG, "NULL context or negative limit");
    HIP_TRY(hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    ctx->hbm_bytes = hbm_bytes > 0 ? std::min<int64_t>(hbm_bytes, (int64_t)tot) : (int64_t)tot;
    return BPX_OK;
}

int bpx_query_footprint(bpx_ctx* ctx, const bpx_db* db, int32_t qlen, const bpx_word_params* params,
                        const bpx_launch* launch, uint32_t* diag_slots, int32_t* waves, size_t* lds_bytes) {
    if (!ctx || !db) return fail(BPX_E_ARG, "NULL handle");
    if (qlen < 0 || qlen > 65535) return fail(BPX_E_ARG, "query length %d outside [0, 65535]", qlen);
    bpx_word_params p{3, 11, 40, 16, 41, 0};
    if (params && params->word_length != 0) p = *params;
    if (p.window <= 0) return fail(BPX_E_ARG, "window must be positive");
    const int64_t D = p.diag_mask > 0 ? (int64_t)p.diag_mask + 1 : next_pow2((int64_t)qlen + p.window);
    if (diag_slots) *diag_slots = (uint32_t)std::min<int64_t>(D, 0xFFFFFFFFll);
    if (waves) *waves = 0;
    if (lds_bytes) *lds_bytes = 0;
    if (D > 65536) return fail(BPX_E_BUDGET, "a %d-residue query needs %lld diagonal slots (at most 65536)", qlen,
                               (long long)D);
    // the planner's view of the query alone, its table read from HBM
    QuerySet set;
    set.ctx = ctx;
    set.nq = 1;
    set.word = p.word_length;
    set.window = p.window;
    set.max_cell.push_back(0);
    bpx::GroupInfo g{};
    g.qres_len = (uint32_t)qlen;
    g.nq = 1;
    g.diag_slots = (uint32_t)D;
    set.groups.push_back(g);
    const Planner P(ctx, set, db->n, count_wide(db, 0, db->n, p.window), launch);
    Dims d;
    d.add(g);
    const int w = P.best_waves(d, false, bpx::kMaxWaves);
    if (w == 0) {
        uint32_t need = 0;
        return P.refuse(d, false, P.fewest(), &need);
    }
    if (waves) *waves = w;
    if (lds_bytes) *lds_bytes = P.plan_of(d, false, w, count_wide(db, 0, db->n, p.window) < db->n).total;
    return BPX_OK;
}

int bpx_scan(bpx_ctx* ctx, const bpx_query* q, const bpx_db* db, uint32_t first, uint32_t count,
             const bpx_launch* launch, uint32_t out_flags, bpx_result** out) {
    if (!q) return fail(BPX_E_ARG, "query is NULL");
    return start_scan(ctx, q->set, db, first, count, launch, out_flags, out);
}

int bpx_scan_batch(bpx_ctx* ctx, const bpx_batch* b, const bpx_db* db, uint32_t first, uint32_t count,
                   const bpx_launch* launch, uint32_t out_flags, bpx_result** out) {
    if (!b) return fail(BPX_E_ARG, "batch is NULL");
    return start_scan(ctx, b->set, db, first, count, launch, out_flags, out);
}

int bpx_scan_noalloc(bpx_ctx* ctx, const bpx_query* q, const bpx_db* db, bpx_init_hsp* d_records,
                     uint32_t* d_subjects, uint64_t record_capacity, bpx_subject_stats* d_stats,
                     uint64_t* num_records, bpx_ungapped_stats* stats) {
    if (!ctx || !q || !db || !num_records) return fail(BPX_E_ARG, "NULL argument");
    *num_records = 0;
    if (record_capacity && (!d_records || !d_subjects)) return fail(BPX_E_ARG, "record buffers are NULL");
    const QuerySet& set = q->set;
    if (set.ctx != ctx || db->ctx != ctx) return fail(BPX_E_ARG, "handles belong to another context");
    if (set.groups.size() != 1 || set.nq != 1) return fail(BPX_E_ARG, "one prepared query expected");
    if (!ctx->scratch) return fail(BPX_E_HIP, "no scratch: %s", ctx->scratch_err);
    if (db->n > BPX_NOALLOC_MAX_SUBJECTS)
        return fail(BPX_E_BUDGET, "%u subjects exceed the allocation-free scan's limit of %u; use bpx_scan", db->n,
                    BPX_NOALLOC_MAX_SUBJECTS);
    if (db->cost_order && set.window > kCostMaxWindow)
        return fail(BPX_E_ARG, "window %d above %d while the database has a cost order", set.window, kCostMaxWindow);
    HIP_TRY(hipSetDevice(ctx->device));
    // the launch plan of bpx_scan for one query: a wide-slot and a
    // narrow-slot class at most, one work item each, in fixed arrays
    const uint32_t count = db->n, nwide = count_wide(db, 0, count, set.window);
    const Planner P(ctx, set, count, nwide, nullptr);
    bool lds = false;
    int waves = 0;
    P.place(set.groups[0], &lds, &waves);
    Dims d;
    d.add(set.groups[0]);
    uint32_t lds_need = 0;
    if (waves == 0) return P.refuse(d, false, P.fewest(), &lds_need);
    bpx_result::Class cls[2];
    bpx::WorkItem items[2] = {};
    uint32_t nc = 0;
    uint64_t nwaves = 0;
    for (int w = 0; w < 2; ++w) {
        const bool narrow = w == 1;
        const uint32_t lo = narrow ? nwide : 0u, hi = narrow ? count : nwide;
        if (lo >= hi) continue;
        cls[nc] = P.make_class(d, lds, false, narrow, waves, lo, hi);
        cls[nc].item_first = nc;
        cls[nc].nitems = 1;
        items[nc] = bpx::WorkItem{0u, lo, hi - lo, 0u};
        nwaves += (uint64_t)ctx->num_cus * P.L.groups_per_cu * waves;
        ++nc;
    }
    const uint64_t seg_cap = nwaves ? BPX_NOALLOC_SEG_RECORDS / nwaves : 1;
    if (seg_cap < 16)
        return fail(BPX_E_BUDGET, "%llu wavefronts leave fewer than 16 record slots each in the scratch",
                    (unsigned long long)nwaves);
    uint8_t* S = ctx->scratch;
    unsigned long long h_meta[kMetaWords + 5] = {};
    ScanJob j;
    j.ctx = ctx;
    j.set = &set;
    j.db = db;
    j.first = 0;
    j.count = count;
    j.classes = cls;
    j.nclasses = nc;
    j.nitems = nc;
    j.groups_per_cu = P.L.groups_per_cu;
    j.sbuf = (uint32_t)P.L.subject_buffer;
    j.stream = ctx->stream;
    j.order = db->d_order;
    j.poff = db->d_poff;
    j.plen = db->d_plen;
    j.cnt = reinterpret_cast<uint32_t*>(S + kNaCnt);
    j.sstats = reinterpret_cast<int32_t*>(d_stats);
    j.run_loc = reinterpret_cast<uint32_t*>(S + kNaRunLoc);
    j.run_len = reinterpret_cast<uint32_t*>(S + kNaRunLen);
    j.seg = reinterpret_cast<bpx::SegRecord*>(S + kNaSeg);
    j.seg_cap = (uint32_t)seg_cap;
    j.tile = reinterpret_cast<uint32_t*>(S + kNaTile);
    j.records = d_records;
    j.subjects = d_subjects;
    j.queries = nullptr;
    j.rec_cap = record_capacity;
    j.meta = reinterpret_cast<unsigned long long*>(S + kNaMeta);
    j.items = reinterpret_cast<bpx::WorkItem*>(S + kNaItems);
    j.work = reinterpret_cast<uint32_t*>(S + kNaWork);
    j.h_meta = h_meta;
    if (nc) HIP_TRY(hipMemcpyAsync(j.items, items, nc * sizeof(bpx::WorkItem), hipMemcpyHostToDevice, ctx->stream));
    const int rc = enqueue_job(j);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    if (e != hipSuccess) return fail(BPX_E_HIP, "allocation-free scan: %s", hipGetErrorString(e));
    const uint32_t ovf = (uint32_t)(h_meta[1] & 0xFFFFFFFFu);
    if (ovf & 2u) return fail(BPX_E_STATE, "internal: hit ring smaller than a lookup cell");
    if (ovf & 1u)
        return fail(BPX_E_BUDGET,
                    "a wavefront saved more than the %llu records its share of the scratch holds; use bpx_scan",
                    (unsigned long long)seg_cap);
    *num_records = h_meta[0];
    if (stats) {
        const unsigned long long* u = h_meta + kMetaWords;
        stats->lookup_hits = (int64_t)u[0];
        stats->num_seqs_lookup_hits = (int64_t)u[1];
        stats->init_extends = (int64_t)u[2];
        stats->good_init_extends = (int64_t)u[3];
        stats->num_seqs_passed = (int64_t)u[4];
    }
    if (h_meta[0] > record_capacity)
        return fail(BPX_E_ARG, "record capacity %llu < %llu records", (unsigned long long)record_capacity,
                    (unsigned long long)h_meta[0]);
    return BPX_OK;
}

int bpx_result_wait(bpx_result* r) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (r->waited) return BPX_OK;
    bpx_ctx* ctx = r->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    for (;;) {
        HIP_TRY(hipStreamSynchronize(r->stream));
        const uint32_t ovf = (uint32_t)(r->h_meta[1] & 0xFFFFFFFFu);
        if (ovf & 2u) return fail(BPX_E_STATE, "internal: hit ring smaller than a lookup cell");
        if (!(ovf & 1u)) break;
        // a wave's record segment filled: rescan with segments as large as
        // the fullest wave needed (remembered for later scans)
        const uint32_t need = (uint32_t)(r->h_meta[2] & 0xFFFFFFFFu);
        const uint32_t cap = std::max<uint32_t>(need + need / 8 + 64, r->seg_cap * 2);
        ctx->seg_hint = std::max(ctx->seg_hint, cap);
        const uint32_t* order = r->d_order ? r->d_order : r->db->d_order;
        result_release(r, true);
        int rc = alloc_buffers(r, cap);
        if (rc) return rc;
        rc = enqueue(r, order, false);
        if (rc) return rc;
        r->graph_stale = r->graph != nullptr;
        ++r->reruns;
    }
    r->total = r->h_meta[0];
    if (r->total) {
        // one query: every query index is 0, nothing to copy
        const bool multi = r->set->nq > 1;
        r->h_records = (bpx_init_hsp*)ctx->pinned.get(r->total * sizeof(bpx_init_hsp));
        r->h_subjects = (uint32_t*)ctx->pinned.get(r->total * 4);
        r->h_queries = multi ? (uint32_t*)ctx->pinned.get(r->total * 4) : nullptr;
        if (!r->h_records || !r->h_subjects || (multi && !r->h_queries))
            return fail(BPX_E_NOMEM, "pinned allocation for records failed");
        HIP_TRY(hipMemcpyAsync(r->h_records, r->d_records, r->total * sizeof(bpx_init_hsp), hipMemcpyDeviceToHost,
                               r->stream));
        HIP_TRY(hipMemcpyAsync(r->h_subjects, r->d_subjects, r->total * 4, hipMemcpyDeviceToHost, r->stream));
        if (multi)
            HIP_TRY(hipMemcpyAsync(r->h_queries, r->d_queries, r->total * 4, hipMemcpyDeviceToHost, r->stream));
    }
    HIP_TRY(hipEventRecord(r->ev[2], r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    HIP_TRY(hipEventElapsedTime(&r->kernel_ms, r->ev[0], r->ev[1]));
    HIP_TRY(hipEventElapsedTime(&r->device_ms, r->ev[0], r->ev[2]));
    r->waited = true;
    return BPX_OK;
}

int bpx_result_ready(const bpx_result* r) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (r->waited) return 1;
    const hipError_t e = hipStreamQuery(r->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(BPX_E_HIP, "stream query: %s", hipGetErrorString(e));
}

int bpx_result_timing(const bpx_result* r, float* kernel_ms, float* device_ms) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (kernel_ms) *kernel_ms = r->kernel_ms;
    if (device_ms) *device_ms = r->device_ms;
    return BPX_OK;
}

int bpx_result_query_stats(const bpx_result* r, bpx_ungapped_stats* stats, uint32_t capacity) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (capacity < r->set->nq) return fail(BPX_E_ARG, "capacity %u < %u queries", capacity, r->set->nq);
    for (uint32_t i = 0; i < r->set->nq && stats; ++i) {
        const unsigned long long* u = r->h_meta + kMetaWords + 5ull * i;
        stats[i].lookup_hits = (int64_t)u[0];
        stats[i].num_seqs_lookup_hits = (int64_t)u[1];
        stats[i].init_extends = (int64_t)u[2];
        stats[i].good_init_extends = (int64_t)u[3];
        stats[i].num_seqs_passed = (int64_t)u[4];
    }
    return BPX_OK;
}

int bpx_result_summary(const bpx_result* r, uint64_t* num_records, bpx_ungapped_stats* stats) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (num_records) *num_records = r->total;
    if (stats) {
        bpx_ungapped_stats s{};
        for (uint32_t i = 0; i < r->set->nq; ++i) {
            const unsigned long long* u = r->h_meta + kMetaWords + 5ull * i;
            s.lookup_hits += (int64_t)u[0];
            s.num_seqs_lookup_hits += (int64_t)u[1];
            s.init_extends += (int64_t)u[2];
            s.good_init_extends += (int64_t)u[3];
            s.num_seqs_passed += (int64_t)u[4];
        }
        *stats = s;
    }
    return BPX_OK;
}

int bpx_result_records(const bpx_result* r, bpx_init_hsp* records, uint32_t* subjects, uint32_t* queries,
                       uint64_t capacity) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (capacity < r->total)
        return fail(BPX_E_ARG, "capacity %llu < %llu records", (unsigned long long)capacity,
                    (unsigned long long)r->total);
    if (records && r->total) std::memcpy(records, r->h_records, r->total * sizeof(bpx_init_hsp));
    if (subjects && r->total) std::memcpy(subjects, r->h_subjects, r->total * 4);
    if (queries && r->total) {
        if (r->h_queries)
            std::memcpy(queries, r->h_queries, r->total * 4);
        else
            std::memset(queries, 0, r->total * 4);
    }
    return BPX_OK;
}

int bpx_result_subject_stats(const bpx_result* r, bpx_subject_stats* stats, uint64_t capacity) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (!(r->flags & BPX_OUT_SUBJECT_STATS)) return fail(BPX_E_STATE, "scan made without BPX_OUT_SUBJECT_STATS");
    const uint64_t pairs = (uint64_t)r->set->nq * r->count;
    if (capacity < pairs)
        return fail(BPX_E_ARG, "capacity %llu < %llu entries", (unsigned long long)capacity, (unsigned long long)pairs);
    if (stats && pairs) std::memcpy(stats, r->h_sstats, pairs * sizeof(bpx_subject_stats));
    return BPX_OK;
}

int bpx_result_replay(bpx_result* r) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!(r->flags & BPX_OUT_GRAPH)) return fail(BPX_E_STATE, "not a graph-mode result (scan without BPX_OUT_GRAPH)");
    if (!r->waited) return fail(BPX_E_STATE, "replay before the previous run was waited");
    bpx_ctx* ctx = r->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    // the previous run's host copies of the records go back to the pool;
    // bpx_result_wait fetches the new ones
    ctx->pinned.put(r->h_records);
    ctx->pinned.put(r->h_subjects);
    ctx->pinned.put(r->h_queries);
    r->h_records = nullptr;
    r->h_subjects = nullptr;
    r->h_queries = nullptr;
    r->waited = false;
    r->total = 0;
    if (r->graph_stale || !r->graph_exec) return graph_capture_and_launch(r);
    HIP_TRY(hipEventRecord(r->ev[0], r->stream));
    HIP_TRY(hipGraphLaunch(r->graph_exec, r->stream));
    HIP_TRY(hipEventRecord(r->ev[1], r->stream));
    return BPX_OK;
}

int bpx_result_phases(const bpx_result* r, bpx_phase_stats* out) {
    if (!r || !out) return fail(BPX_E_ARG, "NULL argument");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (!(r->flags & BPX_OUT_PHASES) || !r->d_phase) return fail(BPX_E_STATE, "scan made without BPX_OUT_PHASES");
    HIP_TRY(hipSetDevice(r->ctx->device));
    const uint64_t nw = total_waves(r);
    std::vector<unsigned long long> h(nw * 8);
    if (nw) HIP_TRY(hipMemcpy(h.data(), r->d_phase, nw * 64, hipMemcpyDeviceToHost));
    constexpr double kTickMs = 1e-5;  // 100 MHz
    bpx_phase_stats st{};
    st.waves = (uint32_t)nw;
    st.classes = (uint32_t)r->classes.size();
    double life = 0, stage = 0, group = 0, flush = 0, life_min = 0, slots = 0;
    uint64_t w0 = 0;
    for (const auto& c : r->classes) {
        const uint64_t cw = (uint64_t)r->ctx->num_cus * r->launch.groups_per_cu * c.waves;
        uint64_t first = ~0ull, last = 0;
        for (uint64_t w = w0; w < w0 + cw; ++w) {
            const unsigned long long* p = &h[w * 8];
            first = std::min<uint64_t>(first, p[0]);
            last = std::max<uint64_t>(last, p[1]);
            const double l = (double)(p[1] - p[0]) * kTickMs;
            life += l;
            life_min = w == 0 ? l : std::min(life_min, l);
            stage += (double)p[2] * kTickMs;
            group += (double)p[3] * kTickMs;
            st.subjects_staged += p[4];
            st.bytes_staged += p[5];
            flush += (double)p[6] * kTickMs;
            st.extensions += p[7];
        }
        if (cw) {
            st.span_ms += (double)(last - first) * kTickMs;
            slots += (double)cw * (double)(last - first) * kTickMs;
        }
        w0 += cw;
    }
    st.busy_fraction = slots > 0 ? life / slots : 0.0;
    if (nw) {
        st.wave_ms_mean = life / (double)nw;
        st.wave_ms_min = life_min;
        st.stage_ms_mean = stage / (double)nw;
        st.group_ms_mean = group / (double)nw;
        st.flush_ms_mean = flush / (double)nw;
    }
    *out = st;
    return BPX_OK;
}

int bpx_result_destroy(bpx_result* r) {
    if (!r) return BPX_OK;
    (void)hipSetDevice(r->ctx->device);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    graph_release(r);
    result_release(r, false);
    for (auto& ev : r->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete r;
    return BPX_OK;
}

}  // extern "C"
