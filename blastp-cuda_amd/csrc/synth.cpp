// synth.cpp -- seeded synthetic queries/databases (see bpx_synth.h).
#include "bpx_synth.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

namespace bpx {
const int8_t* blosum62();
}

namespace {

// ---- random numbers: splitmix64 seeding, xoshiro256** stream -------------
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream, uint64_t index) {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ (index * 0x8CB92BA72F3D8DD7ull);
        for (auto& v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform_open() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }  // (0,1)
    uint32_t below(uint32_t n) { return (uint32_t)(uniform() * n); }
    double normal() {  // Box-Muller, one variate per call
        const double u1 = uniform_open(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    double gamma(double a) {  // Marsaglia & Tsang (2000), boosted for a < 1
        if (a < 1.0) return gamma(a + 1.0) * std::pow(uniform_open(), 1.0 / a);
        const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            const double u = uniform_open();
            if (std::log(u) < 0.5 * x * x + d - d * v + d * std::log(v)) return d * v;
        }
    }
    // geometric on {1,2,...} with the given mean
    uint32_t geometric(double mean) {
        const double p = 1.0 / mean;
        uint32_t k = 1;
        while (uniform() >= p && k < 1000) ++k;
        return k;
    }
};

// Streams: one per kind of draw so that changing one spec field does not
// perturb unrelated bytes.
enum Stream : uint64_t { kQuery = 1, kQLen = 2, kSubject = 3, kPick = 4 };

// ncbistdaa codes of the 20 standard residues and their BLOSUM62 background
// frequencies (renormalised below).
constexpr uint8_t kCodes[20] = {1, 16, 13, 4, 3, 15, 5, 7, 8, 9, 11, 10, 12, 6, 14, 17, 18, 20, 22, 19};
//                              A   R   N  D  C   Q  E  G  H  I   L   K   M  F   P   S   T   W   Y   V
constexpr double kFreq[20] = {.074, .052, .045, .054, .025, .034, .054, .074, .026, .068,
                              .099, .058, .025, .047, .039, .057, .051, .013, .032, .073};

struct Model {
    double cdf[20];
    // conditional substitution CDFs: row a (index 0..19), excluding a itself
    double sub_cdf[20][20];
    Model() {
        double tot = 0;
        for (double f : kFreq) tot += f;
        double acc = 0;
        for (int i = 0; i < 20; ++i) {
            acc += kFreq[i] / tot;
            cdf[i] = acc;
        }
        cdf[19] = 1.0;
        const int8_t* M = bpx::blosum62();
        for (int a = 0; a < 20; ++a) {
            double w[20], s = 0;
            for (int b = 0; b < 20; ++b) {
                w[b] = (a == b) ? 0.0 : (kFreq[b] / tot) * std::pow(2.0, M[kCodes[a] * 28 + kCodes[b]] / 2.0);
                s += w[b];
            }
            double c = 0;
            for (int b = 0; b < 20; ++b) {
                c += w[b] / s;
                sub_cdf[a][b] = c;
            }
            sub_cdf[a][19] = 1.0;
        }
    }
    static int pick(const double* cdf, double u) {
        int i = 0;
        while (i < 19 && u >= cdf[i]) ++i;
        return i;
    }
    uint8_t background(Rng& r) const { return kCodes[pick(cdf, r.uniform())]; }
    uint8_t substitute(Rng& r, uint8_t code) const {
        int a = 0;
        while (a < 20 && kCodes[a] != code) ++a;
        if (a == 20) return background(r);
        return kCodes[pick(sub_cdf[a], r.uniform())];
    }
};

const Model& model() {
    static const Model m;
    return m;
}

uint32_t lognormal_len(Rng& r, double median, double sigma, uint32_t lo, uint32_t hi) {
    const double x = std::exp(std::log(median) + sigma * r.normal());
    double v = std::floor(x + 0.5);
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return (uint32_t)v;
}

// Mutates `src` into a homolog: per position an indel event with
// probability `indel` (insertion or deletion of a geometric run, mean 2),
// otherwise a BLOSUM62-weighted substitution with probability 1-identity.
void make_homolog(Rng& r, const uint8_t* src, uint32_t n, double identity, double indel,
                  std::vector<uint8_t>& out) {
    const Model& m = model();
    out.clear();
    for (uint32_t i = 0; i < n; ++i) {
        if (r.uniform() < indel) {
            const uint32_t k = r.geometric(2.0);
            if (r.uniform() < 0.5) {
                for (uint32_t j = 0; j < k; ++j) out.push_back(m.background(r));
            } else {
                i += k - 1;
                continue;
            }
        }
        out.push_back(r.uniform() < identity ? src[i] : m.substitute(r, src[i]));
    }
}

struct SubjectPlan {
    uint8_t kind;      // 0 background, 1 homolog, 2 biased
    uint32_t source;   // homolog: query index
    uint32_t len;
};

}  // namespace

struct bpx_synth_db {
    std::vector<uint8_t> residues;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> lengths;
    std::vector<uint8_t> kinds;
};

namespace {

// One subject's residues, generated from (seed, index) alone.
void gen_subject(const bpx_synth_db_spec& sp, uint32_t i, const SubjectPlan& plan,
                 const uint8_t* const* queries, const uint32_t* qlens, uint8_t* dst,
                 const std::vector<uint8_t>* homolog) {
    Rng r(sp.seed, kSubject + 200, i);  // content stream, apart from the planning draws
    const Model& m = model();
    if (plan.kind == 2) {
        double p[20], tot = 0;
        for (double& v : p) tot += (v = r.gamma(sp.dirichlet_alpha));
        double cdf[20], acc = 0;
        for (int k = 0; k < 20; ++k) cdf[k] = (acc += p[k] / tot);
        cdf[19] = 1.0;
        for (uint32_t k = 0; k < plan.len; ++k) dst[k] = kCodes[Model::pick(cdf, r.uniform())];
        const uint32_t mlo = std::max<uint32_t>(1, sp.motif_min);
        const uint32_t mhi = std::min<uint32_t>(256, std::max(mlo, sp.motif_max));
        const uint32_t mlen = mlo + r.below(mhi - mlo + 1);
        uint8_t motif[256];
        for (uint32_t k = 0; k < mlen; ++k) motif[k] = kCodes[Model::pick(cdf, r.uniform())];
        const uint32_t rlen = (uint32_t)(plan.len * (0.2 + 0.4 * r.uniform()));
        const uint32_t start = plan.len > rlen ? r.below(plan.len - rlen + 1) : 0;
        for (uint32_t k = 0; k < rlen && start + k < plan.len; ++k) dst[start + k] = motif[k % mlen];
        return;
    }
    for (uint32_t k = 0; k < plan.len; ++k) dst[k] = m.background(r);
    if (plan.kind == 1 && homolog) {
        const uint32_t h = (uint32_t)homolog->size();
        const uint32_t at = plan.len > h ? r.below(plan.len - h + 1) : 0;
        std::memcpy(dst + at, homolog->data(), h);
    }
    (void)queries;
    (void)qlens;
}

}  // namespace

extern "C" {

int bpx_synth_query(uint64_t seed, uint32_t index, uint32_t len, uint8_t* out) {
    if (len && !out) return -1;
    Rng r(seed, kQuery, index);
    const Model& m = model();
    for (uint32_t k = 0; k < len; ++k) out[k] = m.background(r);
    return 0;
}

int bpx_synth_query_lengths(uint64_t seed, uint32_t n, int kind, double lo, double hi,
                            double median, double sigma, uint32_t* out) {
    if ((n && !out) || lo <= 0 || hi < lo) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        Rng r(seed, kQLen, i);
        if (kind == 0) {
            const double x = std::exp(std::log(lo) + r.uniform() * (std::log(hi) - std::log(lo)));
            out[i] = (uint32_t)std::min(hi, std::max(lo, std::floor(x + 0.5)));
        } else {
            out[i] = lognormal_len(r, median, sigma, (uint32_t)lo, (uint32_t)hi);
        }
    }
    return 0;
}

int bpx_synth_db_generate(const bpx_synth_db_spec* spec, const uint8_t* const* queries,
                          const uint32_t* qlens, uint32_t num_queries, int nthreads,
                          bpx_synth_db** out) {
    if (!spec || !out || spec->len_min == 0 || spec->len_max < spec->len_min) return -1;
    const bpx_synth_db_spec sp = *spec;
    const uint32_t n = sp.num_subjects;
    const uint32_t nh = num_queries == 0 ? 0
                        : sp.homolog_count ? std::min(sp.homolog_count, n)
                                           : (uint32_t)std::floor(sp.homolog_fraction * n + 0.5);
    if (nh && (!queries || !qlens)) return -1;

    // Plan: which subjects are homologs (seeded partial shuffle), which are
    // biased, and every background length.
    std::vector<SubjectPlan> plan(n);
    {
        std::vector<uint32_t> idx(n);
        for (uint32_t i = 0; i < n; ++i) idx[i] = i;
        Rng pick(sp.seed, kPick, 0);
        for (uint32_t k = 0; k < nh; ++k) {
            const uint32_t j = k + pick.below(n - k);
            std::swap(idx[k], idx[j]);
            plan[idx[k]].kind = 1;
            plan[idx[k]].source = pick.below(num_queries);
        }
    }
    std::vector<std::vector<uint8_t>> homologs(n);
    for (uint32_t i = 0; i < n; ++i) {
        Rng r(sp.seed, kSubject, i);
        const double u = r.uniform();
        uint32_t len = lognormal_len(r, sp.len_median, sp.len_sigma, sp.len_min, sp.len_max);
        if (plan[i].kind == 1) {
            Rng h(sp.seed, kSubject + 100, i);
            const double ident = sp.identity_min + (sp.identity_max - sp.identity_min) * h.uniform();
            const double indel = sp.indel_min + (sp.indel_max - sp.indel_min) * h.uniform();
            const uint32_t s = plan[i].source;
            make_homolog(h, queries[s], qlens[s], ident, indel, homologs[i]);
            len = std::max<uint32_t>(len, (uint32_t)homologs[i].size());
        } else if (u < sp.biased_fraction) {
            plan[i].kind = 2;
        }
        plan[i].len = len;
    }

    auto* db = new (std::nothrow) bpx_synth_db;
    if (!db) return -1;
    db->lengths.resize(n);
    db->offsets.resize((size_t)n + 1);
    db->kinds.resize(n);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        db->offsets[i] = acc;
        db->lengths[i] = plan[i].len;
        db->kinds[i] = plan[i].kind;
        acc += plan[i].len;
    }
    db->offsets[n] = acc;
    db->residues.resize(acc);

    unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<unsigned>(nt, 64);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            for (uint32_t i = t; i < n; i += nt)
                gen_subject(sp, i, plan[i], queries, qlens, db->residues.data() + db->offsets[i],
                            plan[i].kind == 1 ? &homologs[i] : nullptr);
        });
    for (auto& th : pool) th.join();
    *out = db;
    return 0;
}

int bpx_synth_db_arrays(const bpx_synth_db* db, const uint8_t** residues, const uint64_t** offsets,
                        const uint32_t** lengths, uint32_t* num_subjects, uint64_t* num_residues) {
    if (!db) return -1;
    if (residues) *residues = db->residues.data();
    if (offsets) *offsets = db->offsets.data();
    if (lengths) *lengths = db->lengths.data();
    if (num_subjects) *num_subjects = (uint32_t)db->lengths.size();
    if (num_residues) *num_residues = db->residues.size();
    return 0;
}

int bpx_synth_db_kinds(const bpx_synth_db* db, const uint8_t** kinds) {
    if (!db || !kinds) return -1;
    *kinds = db->kinds.data();
    return 0;
}

int bpx_synth_db_free(bpx_synth_db* db) {
    delete db;
    return 0;
}

}  // extern "C"

// ---- SHA-256 (FIPS 180-4) -------------------------------------------------
namespace {

struct Sha256 {
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint8_t buf[64];
    uint64_t total = 0;
    size_t fill = 0;
    static uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
    void block(const uint8_t* p) {
        static const uint32_t k[64] = {
            0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
            0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
            0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
            0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
            0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
            0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
            0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
            0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const uint32_t s0 = ror(w[i - 15], 7) ^ ror(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = ror(w[i - 2], 17) ^ ror(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int i = 0; i < 64; ++i) {
            const uint32_t t1 = hh + (ror(e, 6) ^ ror(e, 11) ^ ror(e, 25)) + ((e & f) ^ (~e & g)) + k[i] + w[i];
            const uint32_t t2 = (ror(a, 2) ^ ror(a, 13) ^ ror(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
    }
    void update(const void* data, uint64_t len) {
        const uint8_t* p = (const uint8_t*)data;
        total += len;
        while (len) {
            if (fill == 0 && len >= 64) {
                block(p);
                p += 64;
                len -= 64;
                continue;
            }
            const size_t take = std::min<uint64_t>(64 - fill, len);
            std::memcpy(buf + fill, p, take);
            fill += take;
            p += take;
            len -= take;
            if (fill == 64) {
                block(buf);
                fill = 0;
            }
        }
    }
    void hex(char out[65]) {
        const uint64_t bits = total * 8;
        const uint8_t pad = 0x80;
        update(&pad, 1);
        const uint8_t zero = 0;
        while (fill != 56) update(&zero, 1);
        uint8_t lenb[8];
        for (int i = 0; i < 8; ++i) lenb[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(lenb, 8);
        for (int i = 0; i < 8; ++i) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
        out[64] = 0;
    }
};

}  // namespace

extern "C" {

int bpx_sha256_hex(const void* data, uint64_t len, char hex[65]) {
    if (!hex || (len && !data)) return -1;
    Sha256 s;
    s.update(data, len);
    s.hex(hex);
    return 0;
}

int bpx_synth_db_sha256(const bpx_synth_db* db, char hex[65]) {
    if (!db || !hex) return -1;
    Sha256 s;
    for (uint32_t v : db->lengths) {
        const uint8_t le[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        s.update(le, 4);
    }
    s.update(db->residues.data(), db->residues.size());
    s.hex(hex);
    return 0;
}

}  // extern "C"
