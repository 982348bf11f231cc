// runtime.hip -- the C ABI of include/bpx.h: contexts, the resident residue
// arena, query tables, scans and their results.  One HIP stream per
// context; every device buffer is owned by a handle.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "bpx.h"
#include "kernels.hpp"
#include "table.hpp"

namespace bpx {
const int8_t* blosum62();
}

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(BPX_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                               \
    } while (0)

constexpr uint32_t kArenaAlign = 16;   // subject starts in the arena
constexpr uint32_t kArenaTail = 256;   // readable zero bytes after the last subject

// Size-keyed cache of device allocations, so repeated scans do not call
// hipMalloc/hipFree.
struct DevicePool {
    explicit DevicePool(bool host = false) : host_(host) {}
    bool host_;
    std::multimap<size_t, void*> free_;
    ~DevicePool() {
        for (auto& kv : free_) release(kv.second);
    }
    void release(void* p) {
        if (host_)
            (void)hipHostFree(p);
        else
            (void)hipFree(p);
    }
    hipError_t alloc(void** p, size_t bytes) {
        return host_ ? hipHostMalloc(p, bytes, hipHostMallocDefault) : hipMalloc(p, bytes);
    }
    void* get(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            free_.erase(it);
            return p;
        }
        void* p = nullptr;
        if (alloc(&p, bytes) != hipSuccess) {
            // drop cached blocks and retry once
            for (auto& kv : free_) release(kv.second);
            free_.clear();
            if (alloc(&p, bytes) != hipSuccess) return nullptr;
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
    std::map<void*, size_t> sizes_;
};

}  // namespace

struct bpx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    int64_t lds_limit = 0;
    int64_t hbm_bytes = 0;
    uint32_t* work_counter = nullptr;
    DevicePool pool;
    DevicePool pinned{true};
};

struct bpx_db {
    bpx_ctx* ctx = nullptr;
    uint32_t n = 0;
    uint64_t nres = 0;
    uint64_t arena_bytes = 0;
    uint8_t* d_arena = nullptr;
    uint64_t* d_off = nullptr;
    uint32_t* d_len = nullptr;
    uint32_t* d_order = nullptr;        // all subjects, longest first
    std::vector<uint32_t> h_len;
    std::vector<uint32_t> h_order;
};

struct bpx_query {
    bpx_ctx* ctx = nullptr;
    int32_t qlen = 0;
    bpx_word_params params{};
    bpx::HostTable table;
    uint64_t table_bytes = 0;
    uint32_t* d_table = nullptr;
    uint8_t* d_query = nullptr;
    int8_t* d_matrix = nullptr;
};

struct bpx_result {
    bpx_ctx* ctx = nullptr;
    uint32_t first = 0, count = 0;
    uint32_t flags = 0;
    bool waited = false;
    // device
    int32_t* d_stats = nullptr;
    uint32_t* d_rec_loc = nullptr;
    bpx_init_hsp* d_seg = nullptr;
    uint32_t seg_cap = 0, seg_waves = 0;
    uint32_t* d_tile = nullptr;
    uint32_t* d_rec_off = nullptr;
    bpx_init_hsp* d_records = nullptr;
    uint32_t* d_subjects = nullptr;
    unsigned long long* d_ustats = nullptr;  // 5 stats + total + overflow word
    uint32_t* d_overflow = nullptr;
    uint32_t* d_order = nullptr;             // owned only for sub-range scans
    // host
    unsigned long long* h_ustats = nullptr;  // pinned: 5 stats, total, overflow
    bpx_init_hsp* h_records = nullptr;     // pinned
    uint32_t* h_subjects = nullptr;        // pinned
    bpx_subject_stats* h_sstats = nullptr; // pinned
    bpx_ungapped_stats ustats{};
    uint64_t total = 0;
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float kernel_ms = 0.f, device_ms = 0.f;
    // relaunch state (segment overflow)
    const bpx_query* q = nullptr;
    const bpx_db* db = nullptr;
    bpx_launch launch{};
    int reruns = 0;
};

namespace {

void result_release(bpx_result* r) {
    auto& pool = r->ctx->pool;
    pool.put(r->d_stats);
    pool.put(r->d_rec_loc);
    pool.put(r->d_seg);
    pool.put(r->d_tile);
    pool.put(r->d_rec_off);
    pool.put(r->d_records);
    pool.put(r->d_subjects);
    pool.put(r->d_ustats);
    pool.put(r->d_order);
    r->ctx->pinned.put(r->h_records);
    r->ctx->pinned.put(r->h_subjects);
    r->ctx->pinned.put(r->h_sstats);
    r->h_records = nullptr;
    r->h_subjects = nullptr;
    r->h_sstats = nullptr;
    r->d_stats = nullptr;
    r->d_rec_loc = nullptr;
    r->d_seg = nullptr;
    r->d_tile = nullptr;
    r->d_rec_off = nullptr;
    r->d_records = nullptr;
    r->d_subjects = nullptr;
    r->d_ustats = nullptr;
    r->d_overflow = nullptr;
    r->d_order = nullptr;
}

bpx_launch resolve_launch(const bpx_launch* in, const bpx::HostTable& t, int32_t qlen = 0,
                          int32_t diag_len = 0, int64_t lds_limit = 0) {
    bpx_launch l{};
    if (in) l = *in;
    const bool auto_waves = l.waves_per_group <= 0;
    if (auto_waves) l.waves_per_group = 8;
    l.waves_per_group = std::min(16, std::max(1, l.waves_per_group));
    if (l.groups_per_cu <= 0) l.groups_per_cu = 1;
    if (l.subjects_per_grab <= 0) l.subjects_per_grab = 1;
    uint32_t qc = l.queue_capacity > 0 ? (uint32_t)l.queue_capacity : 256u;
    const uint32_t need = std::max<uint32_t>(128u, t.max_cell + 64u);
    while (qc < need) qc <<= 1;
    uint32_t p2 = 1;
    while (p2 < qc) p2 <<= 1;
    l.queue_capacity = (int32_t)p2;
    // Long queries have large diagonal tables: with the default shape, use
    // fewer wavefronts per workgroup until the per-wave LDS fits.
    if (auto_waves && lds_limit > 0)
        while (l.waves_per_group > 1 &&
               bpx::plan_lds((uint32_t)t.packed_bytes(), qlen, diag_len, p2, l.waves_per_group,
                             (uint32_t)lds_limit, true).total > (uint64_t)lds_limit)
            l.waves_per_group /= 2;
    return l;
}

int32_t diag_len_of(const bpx_word_params& p) { return p.diag_mask + 1; }

// Enqueues the word finder and the compaction for result r (buffers already
// allocated).
int enqueue_scan(bpx_result* r, const bpx_query* q, const bpx_db* db, const bpx_launch& L,
                 const bpx::LdsPlan& plan, const uint32_t* order) {
    bpx_ctx* ctx = r->ctx;
    HIP_TRY(bpx::launch_fill_u32(reinterpret_cast<uint32_t*>(r->d_ustats), 0u, 16, ctx->stream));
    HIP_TRY(bpx::launch_fill_u32(ctx->work_counter, 0u, 1, ctx->stream));

    bpx::ScanParams sp{};
    sp.residues = db->d_arena;
    sp.soff = db->d_off;
    sp.slen = db->d_len;
    sp.order = order;
    sp.norder = r->count;
    sp.first = r->first;
    sp.table = q->d_table;
    sp.table_words = (uint32_t)(q->table_bytes / 4);
    sp.ngroups = q->table.ngroups;
    sp.npresent = q->table.npresent;
    sp.query = q->d_query;
    sp.qlen = q->qlen;
    sp.matrix = q->d_matrix;
    sp.word = q->params.word_length;
    sp.window = q->params.window;
    sp.xdrop = q->params.x_dropoff;
    sp.cutoff = q->params.cutoff_score;
    sp.diag_mask = q->params.diag_mask;
    sp.work_counter = ctx->work_counter;
    sp.grab = (uint32_t)L.subjects_per_grab;
    sp.stats = r->d_stats;
    sp.rec_loc = r->d_rec_loc;
    sp.seg = r->d_seg;
    sp.seg_cap = r->seg_cap;
    sp.overflow = r->d_overflow;
    sp.lds_table = plan.table;
    sp.lds_matrix = plan.matrix;
    sp.lds_query = plan.query;
    sp.lds_wave = plan.wave;
    sp.wave_stride = plan.wave_stride;
    sp.qcap = (uint32_t)L.queue_capacity;
    sp.lds_total = plan.total;

    const int groups = ctx->num_cus * L.groups_per_cu;
    HIP_TRY(hipEventRecord(r->ev[0], ctx->stream));
    if (r->count > 0) HIP_TRY(bpx::launch_scan(sp, plan.table_in_lds, L.waves_per_group, groups, ctx->stream));
    HIP_TRY(hipEventRecord(r->ev[1], ctx->stream));

    bpx::CompactParams cp{};
    cp.stats = r->d_stats;
    cp.rec_loc = r->d_rec_loc;
    cp.seg = r->d_seg;
    cp.count = r->count;
    cp.first = r->first;
    cp.tile_sum = r->d_tile;
    cp.rec_off = r->d_rec_off;
    cp.records = r->d_records;
    cp.subjects = r->d_subjects;
    cp.ustats = r->d_ustats;
    cp.overflow = r->d_overflow;
    HIP_TRY(bpx::launch_compact(cp, ctx->stream));
    HIP_TRY(hipMemcpyAsync(r->h_ustats, r->d_ustats, 7 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, ctx->stream));
    if ((r->flags & BPX_OUT_SUBJECT_STATS) && r->count) {
        if (!r->h_sstats)
            r->h_sstats = (bpx_subject_stats*)ctx->pinned.get((size_t)r->count * sizeof(bpx_subject_stats));
        if (!r->h_sstats) return fail(BPX_E_NOMEM, "pinned allocation for subject stats failed");
        HIP_TRY(hipMemcpyAsync(r->h_sstats, r->d_stats, (size_t)r->count * sizeof(bpx_subject_stats),
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    return BPX_OK;
}

int alloc_result_buffers(bpx_result* r, uint32_t seg_cap, uint32_t seg_waves) {
    auto& pool = r->ctx->pool;
    const uint32_t cnt = std::max<uint32_t>(r->count, 1);
    const uint32_t ntiles = (cnt + bpx::kTile - 1) / bpx::kTile;
    const uint64_t seg_total = (uint64_t)seg_cap * seg_waves;
    r->seg_cap = seg_cap;
    r->seg_waves = seg_waves;
    r->d_stats = (int32_t*)pool.get((size_t)cnt * 3 * sizeof(int32_t));
    r->d_rec_loc = (uint32_t*)pool.get((size_t)cnt * sizeof(uint32_t));
    r->d_seg = (bpx_init_hsp*)pool.get(seg_total * sizeof(bpx_init_hsp));
    r->d_tile = (uint32_t*)pool.get((size_t)ntiles * sizeof(uint32_t));
    r->d_rec_off = (uint32_t*)pool.get(((size_t)cnt + 1) * sizeof(uint32_t));
    r->d_records = (bpx_init_hsp*)pool.get(seg_total * sizeof(bpx_init_hsp));
    r->d_subjects = (uint32_t*)pool.get(seg_total * sizeof(uint32_t));
    r->d_ustats = (unsigned long long*)pool.get(8 * sizeof(unsigned long long));
    if (!r->d_stats || !r->d_rec_loc || !r->d_seg || !r->d_tile || !r->d_rec_off || !r->d_records ||
        !r->d_subjects || !r->d_ustats)
        return fail(BPX_E_NOMEM, "device allocation for scan results failed (%llu segment records)",
                    (unsigned long long)seg_total);
    r->d_overflow = reinterpret_cast<uint32_t*>(r->d_ustats + 6);
    return BPX_OK;
}

}  // namespace

extern "C" {

const char* bpx_last_error(void) { return g_err.c_str(); }

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
        hipMalloc(&ctx->work_counter, 256) != hipSuccess) {
        delete ctx;
        return fail(BPX_E_HIP, "stream / counter allocation failed");
    }
    *out = ctx;
    return BPX_OK;
}

int bpx_context_destroy(bpx_ctx* ctx) {
    if (!ctx) return BPX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->work_counter);
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
    // Repack into the device layout: 16-byte-aligned subject starts, zero
    // padding, and a readable zero tail.
    std::vector<uint64_t> off(num_subjects);
    uint64_t at = 0, nres = 0;
    for (uint32_t i = 0; i < num_subjects; ++i) {
        if (lengths[i] > 0x7FFFFFFFu) return fail(BPX_E_ARG, "subject %u too long", i);
        off[i] = at;
        at += ((uint64_t)lengths[i] + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        nres += lengths[i];
    }
    const uint64_t arena = at + kArenaTail;
    std::vector<uint8_t> host(arena, 0);
    for (uint32_t i = 0; i < num_subjects; ++i) {
        const uint8_t* src = residues + offsets[i];
        for (uint32_t k = 0; k < lengths[i]; ++k)
            if (src[k] >= BPX_ALPHABET)
                return fail(BPX_E_ARG, "subject %u position %u: residue code %u >= 28", i, k, src[k]);
        std::memcpy(host.data() + off[i], src, lengths[i]);
    }
    auto* db = new (std::nothrow) bpx_db;
    if (!db) return fail(BPX_E_NOMEM, "host allocation failed");
    db->ctx = ctx;
    db->n = num_subjects;
    db->nres = nres;
    db->arena_bytes = arena;
    db->h_len.assign(lengths, lengths + num_subjects);
    db->h_order.resize(num_subjects);
    std::iota(db->h_order.begin(), db->h_order.end(), 0u);
    // longest first (ties by index): the scan's dynamic queue then behaves
    // like longest-processing-time-first list scheduling
    std::stable_sort(db->h_order.begin(), db->h_order.end(),
                     [&](uint32_t a, uint32_t b) { return db->h_len[a] > db->h_len[b]; });
    const size_t n1 = std::max<uint32_t>(num_subjects, 1);
    if (hipMalloc(&db->d_arena, arena) != hipSuccess || hipMalloc(&db->d_off, n1 * 8) != hipSuccess ||
        hipMalloc(&db->d_len, n1 * 4) != hipSuccess || hipMalloc(&db->d_order, n1 * 4) != hipSuccess) {
        bpx_db_destroy(db);
        return fail(BPX_E_NOMEM, "device allocation of %llu-byte arena failed", (unsigned long long)arena);
    }
    hipError_t e = hipMemcpy(db->d_arena, host.data(), arena, hipMemcpyHostToDevice);
    if (e == hipSuccess && num_subjects) {
        e = hipMemcpy(db->d_off, off.data(), (size_t)num_subjects * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(db->d_len, lengths, (size_t)num_subjects * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(db->d_order, db->h_order.data(), (size_t)num_subjects * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        bpx_db_destroy(db);
        return fail(BPX_E_HIP, "arena upload failed: %s", hipGetErrorString(e));
    }
    *out = db;
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
    if (p.diag_mask == 0) {
        int32_t n = 1;
        while (n < qlen + p.window) n <<= 1;
        p.diag_mask = n - 1;
    }
    if (p.diag_mask < 0 || ((p.diag_mask + 1) & p.diag_mask) != 0 || p.diag_mask + 1 > (1 << 16))
        return fail(BPX_E_ARG, "diag_mask must be 2^k - 1 with 2^k <= 65536");
    const int8_t* M = matrix ? matrix : bpx::blosum62();
    HIP_TRY(hipSetDevice(ctx->device));
    auto* q = new (std::nothrow) bpx_query;
    if (!q) return fail(BPX_E_NOMEM, "host allocation failed");
    q->ctx = ctx;
    q->qlen = qlen;
    q->params = p;
    if (!bpx::build_table(query, qlen, M, p.threshold, p.word_length, &q->table)) {
        delete q;
        return fail(BPX_E_ARG, "lookup table build failed (residue codes must be < 28)");
    }
    const std::vector<uint32_t> blob = q->table.pack();
    q->table_bytes = (uint64_t)blob.size() * 4;
    const size_t qbytes = std::max<size_t>((size_t)qlen, 1);
    if (hipMalloc(&q->d_table, q->table_bytes) != hipSuccess || hipMalloc(&q->d_query, qbytes) != hipSuccess ||
        hipMalloc(&q->d_matrix, 28 * 28) != hipSuccess) {
        bpx_query_destroy(q);
        return fail(BPX_E_NOMEM, "device allocation for the query table failed");
    }
    hipError_t e = hipMemcpy(q->d_table, blob.data(), q->table_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && qlen) e = hipMemcpy(q->d_query, query, (size_t)qlen, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(q->d_matrix, M, 28 * 28, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        bpx_query_destroy(q);
        return fail(BPX_E_HIP, "query upload failed: %s", hipGetErrorString(e));
    }
    *out = q;
    return BPX_OK;
}

int bpx_query_info(const bpx_query* q, int64_t* table_bytes, int32_t* lds_resident, int32_t* num_present_words,
                   int64_t* num_entries) {
    if (!q) return fail(BPX_E_ARG, "query is NULL");
    if (table_bytes) *table_bytes = (int64_t)q->table_bytes;
    if (lds_resident) {
        const bpx_launch L = resolve_launch(nullptr, q->table, q->qlen, diag_len_of(q->params), q->ctx->lds_limit);
        const bpx::LdsPlan pl = bpx::plan_lds((uint32_t)q->table_bytes, q->qlen, diag_len_of(q->params),
                                              (uint32_t)L.queue_capacity, L.waves_per_group,
                                              (uint32_t)q->ctx->lds_limit, true);
        *lds_resident = pl.table_in_lds ? 1 : 0;
    }
    if (num_present_words) *num_present_words = (int32_t)q->table.npresent;
    if (num_entries) *num_entries = (int64_t)q->table.entries.size();
    return BPX_OK;
}

int bpx_query_destroy(bpx_query* q) {
    if (!q) return BPX_OK;
    (void)hipFree(q->d_table);
    (void)hipFree(q->d_query);
    (void)hipFree(q->d_matrix);
    delete q;
    return BPX_OK;
}

int bpx_check_memory(bpx_ctx* ctx, const bpx_db* db, const bpx_query* q, uint32_t first, uint32_t count,
                     const bpx_launch* launch, size_t* device_bytes, size_t* lds_bytes) {
    if (!ctx || !db || !q) return fail(BPX_E_ARG, "NULL handle");
    if ((uint64_t)first + count > db->n) return fail(BPX_E_ARG, "subject range outside the database");
    const bpx_launch L = resolve_launch(launch, q->table, q->qlen, diag_len_of(q->params), ctx->lds_limit);
    const bpx::LdsPlan pl = bpx::plan_lds((uint32_t)q->table_bytes, q->qlen, diag_len_of(q->params),
                                          (uint32_t)L.queue_capacity, L.waves_per_group,
                                          (uint32_t)ctx->lds_limit, true);
    const uint64_t waves = (uint64_t)ctx->num_cus * L.groups_per_cu * L.waves_per_group;
    const uint64_t seg_cap = std::max<uint64_t>(64, ((uint64_t)count / 2 + 65536) / waves + 1);
    const uint64_t need = db->arena_bytes + (uint64_t)db->n * 16 + q->table_bytes + (uint64_t)count * 24 +
                          seg_cap * waves * (24 * 2 + 4);
    if (device_bytes) *device_bytes = (size_t)need;
    if (lds_bytes) *lds_bytes = pl.total;
    if (pl.total > (uint64_t)ctx->lds_limit)
        return fail(BPX_E_BUDGET, "LDS plan %u B exceeds %lld B", pl.total, (long long)ctx->lds_limit);
    if (ctx->hbm_bytes && need > (uint64_t)ctx->hbm_bytes)
        return fail(BPX_E_BUDGET, "needs %llu device bytes", (unsigned long long)need);
    return BPX_OK;
}

int bpx_scan(bpx_ctx* ctx, const bpx_query* q, const bpx_db* db, uint32_t first, uint32_t count,
             const bpx_launch* launch, uint32_t out_flags, bpx_result** out) {
    if (!ctx || !q || !db || !out) return fail(BPX_E_ARG, "NULL handle");
    *out = nullptr;
    if (q->ctx != ctx || db->ctx != ctx) return fail(BPX_E_ARG, "handles belong to another context");
    if ((uint64_t)first + count > db->n) return fail(BPX_E_ARG, "subject range outside the database");
    HIP_TRY(hipSetDevice(ctx->device));
    const bpx_launch L = resolve_launch(launch, q->table, q->qlen, diag_len_of(q->params), ctx->lds_limit);
    const bpx::LdsPlan plan = bpx::plan_lds((uint32_t)q->table_bytes, q->qlen, diag_len_of(q->params),
                                            (uint32_t)L.queue_capacity, L.waves_per_group,
                                            (uint32_t)ctx->lds_limit, true);
    if (plan.total > (uint64_t)ctx->lds_limit)
        return fail(BPX_E_BUDGET, "LDS plan %u B exceeds the %lld B per workgroup (diag %d, %d waves)",
                    plan.total, (long long)ctx->lds_limit, diag_len_of(q->params), L.waves_per_group);

    auto* r = new (std::nothrow) bpx_result;
    if (!r) return fail(BPX_E_NOMEM, "host allocation failed");
    r->ctx = ctx;
    r->first = first;
    r->count = count;
    r->flags = out_flags;
    r->q = q;
    r->db = db;
    r->launch = L;
    for (auto& ev : r->ev)
        if (hipEventCreate(&ev) != hipSuccess) {
            bpx_result_destroy(r);
            return fail(BPX_E_HIP, "event creation failed");
        }
    if (hipHostMalloc((void**)&r->h_ustats, 8 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        bpx_result_destroy(r);
        return fail(BPX_E_NOMEM, "pinned allocation failed");
    }
    const uint32_t waves = (uint32_t)(ctx->num_cus * L.groups_per_cu * L.waves_per_group);
    const uint32_t seg_cap = std::max<uint32_t>(64, (uint32_t)(((uint64_t)count / 2 + 65536) / waves + 1));
    int rc = alloc_result_buffers(r, seg_cap, waves);
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
        r->d_order = (uint32_t*)ctx->pool.get(std::max<size_t>(count, 1) * 4);
        if (!r->d_order) {
            bpx_result_destroy(r);
            return fail(BPX_E_NOMEM, "work list allocation failed");
        }
        if (count) {
            hipError_t e = hipMemcpyAsync(r->d_order, sub.data(), (size_t)count * 4, hipMemcpyHostToDevice,
                                          ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                bpx_result_destroy(r);
                return fail(BPX_E_HIP, "work list upload failed: %s", hipGetErrorString(e));
            }
        }
        order = r->d_order;
    }
    rc = enqueue_scan(r, q, db, L, plan, order);
    if (rc) {
        bpx_result_destroy(r);
        return rc;
    }
    *out = r;
    return BPX_OK;
}

int bpx_result_wait(bpx_result* r) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (r->waited) return BPX_OK;
    bpx_ctx* ctx = r->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    for (;;) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        const uint32_t ovf = (uint32_t)(r->h_ustats[6] & 0xFFFFFFFFu);
        if (ovf & 2u) return fail(BPX_E_STATE, "internal: hit ring smaller than a lookup cell");
        if (!(ovf & 1u)) break;
        // a wave's record segment filled: rescan with 4x the segment size
        const uint32_t cap = r->seg_cap * 4, waves = r->seg_waves;
        const uint32_t* order = r->d_order ? r->d_order : r->db->d_order;
        uint32_t* keep_order = r->d_order;
        r->d_order = nullptr;
        result_release(r);
        r->d_order = keep_order;
        int rc = alloc_result_buffers(r, cap, waves);
        if (rc) return rc;
        const bpx::LdsPlan plan = bpx::plan_lds((uint32_t)r->q->table_bytes, r->q->qlen,
                                                diag_len_of(r->q->params), (uint32_t)r->launch.queue_capacity,
                                                r->launch.waves_per_group, (uint32_t)ctx->lds_limit, true);
        rc = enqueue_scan(r, r->q, r->db, r->launch, plan, order);
        if (rc) return rc;
        ++r->reruns;
    }
    r->total = r->h_ustats[5];
    r->ustats.lookup_hits = (int64_t)r->h_ustats[0];
    r->ustats.num_seqs_lookup_hits = (int64_t)r->h_ustats[1];
    r->ustats.init_extends = (int64_t)r->h_ustats[2];
    r->ustats.good_init_extends = (int64_t)r->h_ustats[3];
    r->ustats.num_seqs_passed = (int64_t)r->h_ustats[4];
    if (r->total) {
        r->h_records = (bpx_init_hsp*)ctx->pinned.get(r->total * sizeof(bpx_init_hsp));
        r->h_subjects = (uint32_t*)ctx->pinned.get(r->total * sizeof(uint32_t));
        if (!r->h_records || !r->h_subjects) return fail(BPX_E_NOMEM, "pinned allocation for records failed");
        HIP_TRY(hipMemcpyAsync(r->h_records, r->d_records, r->total * sizeof(bpx_init_hsp),
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(r->h_subjects, r->d_subjects, r->total * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipEventRecord(r->ev[2], ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipEventElapsedTime(&r->kernel_ms, r->ev[0], r->ev[1]));
    HIP_TRY(hipEventElapsedTime(&r->device_ms, r->ev[0], r->ev[2]));
    r->waited = true;
    return BPX_OK;
}

int bpx_result_timing(const bpx_result* r, float* kernel_ms, float* device_ms) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (kernel_ms) *kernel_ms = r->kernel_ms;
    if (device_ms) *device_ms = r->device_ms;
    return BPX_OK;
}

int bpx_result_summary(const bpx_result* r, uint64_t* num_records, bpx_ungapped_stats* stats) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (num_records) *num_records = r->total;
    if (stats) *stats = r->ustats;
    return BPX_OK;
}

int bpx_result_records(const bpx_result* r, bpx_init_hsp* records, uint32_t* subjects, uint64_t capacity) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (capacity < r->total) return fail(BPX_E_ARG, "capacity %llu < %llu records", (unsigned long long)capacity,
                                         (unsigned long long)r->total);
    if (records && r->total) std::memcpy(records, r->h_records, r->total * sizeof(bpx_init_hsp));
    if (subjects && r->total) std::memcpy(subjects, r->h_subjects, r->total * sizeof(uint32_t));
    return BPX_OK;
}

int bpx_result_subject_stats(const bpx_result* r, bpx_subject_stats* stats, uint32_t capacity) {
    if (!r) return fail(BPX_E_ARG, "result is NULL");
    if (!r->waited) return fail(BPX_E_STATE, "result not waited");
    if (!(r->flags & BPX_OUT_SUBJECT_STATS)) return fail(BPX_E_STATE, "scan made without BPX_OUT_SUBJECT_STATS");
    if (capacity < r->count) return fail(BPX_E_ARG, "capacity %u < %u subjects", capacity, r->count);
    if (stats && r->count) std::memcpy(stats, r->h_sstats, (size_t)r->count * sizeof(bpx_subject_stats));
    return BPX_OK;
}

int bpx_result_destroy(bpx_result* r) {
    if (!r) return BPX_OK;
    (void)hipSetDevice(r->ctx->device);
    (void)hipStreamSynchronize(r->ctx->stream);
    result_release(r);
    if (r->h_ustats) (void)hipHostFree(r->h_ustats);
    for (auto& ev : r->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete r;
    return BPX_OK;
}

}  // extern "C"
