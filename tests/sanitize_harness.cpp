// sanitize_harness.cpp -- drives the host-only library sources (lookup-table
// builder, synthetic generator, NCBI volume and .gpu readers) and the CPU
// oracle through their C interfaces, for builds under AddressSanitizer +
// UndefinedBehaviorSanitizer and under ThreadSanitizer (the generator and the
// oracle's whole-database run are threaded).  tests/test_sanitizers.py builds
// and runs it; it exits non-zero on any unexpected result.
//
// usage: sanitize_harness <scratch directory>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "bpx.h"
#include "bpx_synth.h"
#include "oracle.h"

namespace bpx {
const int8_t* blosum62();  // the library's BLOSUM62 table (blosum62.cpp)
}

// The device loader is not part of this host-only build; the volume reader's
// loading entry point (bpx_db_load_volumes) is not called here.
extern "C" int bpx_db_load(bpx_ctx*, const uint8_t*, const uint64_t*, const uint32_t*, uint32_t, bpx_db**) {
    return BPX_E_NODEV;
}

static int g_failures = 0;
#define CHECK(cond)                                                                \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            ++g_failures;                                                          \
        }                                                                          \
    } while (0)

static std::vector<uint8_t> read_all(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

static void write_all(const std::string& path, const std::vector<uint8_t>& b) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(b.data()), (std::streamsize)b.size());
}

static void put_be32(std::vector<uint8_t>& b, size_t at, uint32_t v) {
    b[at] = (uint8_t)(v >> 24);
    b[at + 1] = (uint8_t)(v >> 16);
    b[at + 2] = (uint8_t)(v >> 8);
    b[at + 3] = (uint8_t)v;
}

static uint32_t get_be32(const std::vector<uint8_t>& b, size_t at) {
    return (uint32_t)b[at] << 24 | (uint32_t)b[at + 1] << 16 | (uint32_t)b[at + 2] << 8 | b[at + 3];
}

// the library's table against the oracle's brute-force one
static void tables() {
    const int8_t* M = bpx::blosum62();
    for (int word = 2; word <= 3; ++word)
        for (uint32_t qlen : {0u, 1u, 2u, 3u, 50u, 700u}) {
            std::vector<uint8_t> q(qlen + 1);
            CHECK(bpx_synth_query(11, qlen, qlen, q.data()) == BPX_OK);
            const uint32_t nwords = 1u << (5 * word);
            std::vector<uint32_t> starts(nwords + 1);
            uint64_t n = 0;
            std::vector<uint16_t> entries(1);
            const int rc = bpx_table_build(q.data(), (int32_t)qlen, M, 11, word, starts.data(), entries.data(), 0, &n);
            CHECK(rc == BPX_OK || rc == BPX_E_ARG);
            entries.resize(n + 1);
            CHECK(bpx_table_build(q.data(), (int32_t)qlen, M, 11, word, starts.data(), entries.data(), n, &n) ==
                  BPX_OK);
            or_lookup lk{};
            CHECK(or_lookup_build(q.data(), (int32_t)qlen, M, 11, word, &lk) == 0);
            CHECK(lk.cell_start[nwords] == n);
            for (uint32_t w = 0; w <= nwords; ++w) CHECK(lk.cell_start[w] == starts[w]);
            for (uint64_t e = 0; e < n; ++e) CHECK(lk.entries[e] == entries[e]);
            or_lookup_free(&lk);
        }
    std::vector<uint8_t> bad = {1, 2, 40};
    uint32_t starts[1025];
    uint16_t e[4];
    uint64_t n = 0;
    CHECK(bpx_table_build(bad.data(), 3, M, 11, 2, starts, e, 4, &n) != BPX_OK);
}

// a threaded synthetic database, the oracle's threaded run over it, and the
// volume files written from it
static void database_and_volumes(const std::string& dir) {
    std::vector<uint8_t> q(400);
    CHECK(bpx_synth_query(12, 0, 400, q.data()) == BPX_OK);
    const uint8_t* qs[1] = {q.data()};
    const uint32_t ql[1] = {400};
    bpx_synth_db_spec spec{};
    spec.seed = 13;
    spec.num_subjects = 3000;
    spec.len_median = 300;
    spec.len_sigma = 0.8;
    spec.len_min = 20;
    spec.len_max = 6000;
    spec.homolog_fraction = 0.03;
    spec.identity_min = 0.4;
    spec.identity_max = 0.7;
    spec.indel_max = 0.03;
    spec.biased_fraction = 0.2;
    spec.dirichlet_alpha = 0.3;
    spec.motif_min = 5;
    spec.motif_max = 30;
    bpx_synth_db* sdb = nullptr;
    CHECK(bpx_synth_db_generate(&spec, qs, ql, 1, 4, &sdb) == BPX_OK);
    const uint8_t* res = nullptr;
    const uint64_t* off = nullptr;
    const uint32_t* len = nullptr;
    uint32_t n = 0;
    uint64_t nres = 0;
    CHECK(bpx_synth_db_arrays(sdb, &res, &off, &len, &n, &nres) == BPX_OK);
    char hex[65];
    CHECK(bpx_synth_db_sha256(sdb, hex) == BPX_OK && std::strlen(hex) == 64);

    // the oracle, threaded in 4,096-subject chunks: one thread and four agree
    or_lookup lk{};
    const int8_t* M = or_blosum62();
    CHECK(or_lookup_build(q.data(), 400, M, 11, 3, &lk) == 0);
    or_params p{3, 11, 40, 16, 41, or_diag_length(400, 40) - 1};
    std::vector<int32_t> c1(3 * (size_t)n), c4(3 * (size_t)n);
    std::vector<uint64_t> ro1(n + 1), ro4(n + 1);
    or_hsp *r1 = nullptr, *r4 = nullptr;
    CHECK(or_run_db(&lk, q.data(), 400, M, &p, res, off, len, n, 1, c1.data(), ro1.data(), &r1) == 0);
    CHECK(or_run_db(&lk, q.data(), 400, M, &p, res, off, len, n, 4, c4.data(), ro4.data(), &r4) == 0);
    CHECK(c1 == c4 && ro1 == ro4);
    CHECK(std::memcmp(r1, r4, ro1[n] * sizeof(or_hsp)) == 0);
    CHECK(ro1[n] > 0);
    or_free(r1);
    or_free(r4);
    or_lookup_free(&lk);

    // volume round trip
    const std::string base = dir + "/v";
    CHECK(bpx_volume_write(base.c_str(), "harness", res, off, len, n) == BPX_OK);
    bpx_volume* v = nullptr;
    CHECK(bpx_volume_read(base.c_str(), &v) == BPX_OK);
    const uint8_t* vres = nullptr;
    const uint64_t* voff = nullptr;
    const uint32_t* vlen = nullptr;
    CHECK(bpx_volume_arrays(v, &vres, &voff, &vlen) == BPX_OK);
    CHECK(std::memcmp(vres, res, nres) == 0 && std::memcmp(vlen, len, 4ull * n) == 0);
    bpx_volume_destroy(v);
    bpx_synth_db_free(sdb);

    // hostile index headers: a sequence count of 2^32 - 1, a first sequence
    // offset of 2^32 - 1, and a file cut short
    const std::vector<uint8_t> pin = read_all(base + ".pin");
    const uint32_t tlen = get_be32(pin, 8);
    const uint32_t dlen = get_be32(pin, 12 + tlen);
    const size_t at_n = 16 + tlen + dlen;
    const size_t at_so = at_n + 4 + 8 + 4 + 4ull * (n + 1);
    const std::vector<uint8_t> psq = read_all(base + ".psq");
    auto expect_refused = [&](const std::vector<uint8_t>& bad, const char* name) {
        const std::string b = dir + "/" + name;
        write_all(b + ".pin", bad);
        write_all(b + ".psq", psq);
        bpx_volume* h = nullptr;
        CHECK(bpx_volume_read(b.c_str(), &h) == BPX_E_ARG && h == nullptr);
    };
    std::vector<uint8_t> huge = pin;
    put_be32(huge, at_n, 0xFFFFFFFFu);
    expect_refused(huge, "huge");
    std::vector<uint8_t> wrap = pin;
    put_be32(wrap, at_so, 0xFFFFFFFFu);
    expect_refused(wrap, "wrap");
    expect_refused(std::vector<uint8_t>(pin.begin(), pin.begin() + (ptrdiff_t)(pin.size() / 2)), "cut");

    // alias lists and malformed .gpuinfo text
    std::ofstream(dir + "/db.pal") << "TITLE t\nDBLIST v\n";
    char names[4096];
    size_t need = 0;
    uint32_t count = 0;
    CHECK(bpx_db_volumes((dir + "/db").c_str(), names, sizeof names, &need, &count) == BPX_OK && count == 1);
    std::ofstream(dir + "/bad.gpuinfo") << "#Number of blocksx\nnot a number\n";
    uint32_t nv = 0, gs = 0, pct = 0;
    CHECK(bpx_gpu_info((dir + "/bad.gpuinfo").c_str(), &nv, &gs, &pct) != BPX_OK);
    bpx_volume* gv = nullptr;
    CHECK(bpx_gpu_volume_read((dir + "/bad.gpuinfo").c_str(), 0, &gv) != BPX_OK && gv == nullptr);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
        return 2;
    }
    tables();
    database_and_volumes(argv[1]);
    if (g_failures) {
        std::fprintf(stderr, "%d checks failed\n", g_failures);
        return 1;
    }
    std::printf("sanitized host run: all checks passed\n");
    return 0;
}
