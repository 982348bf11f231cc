// volume.cpp -- NCBI BLAST protein database volumes (format version 4) into
// the resident residue arena.
//
//   .pin  index, big-endian 32-bit fields unless noted:
//         version (4), database type (1 = protein), title length + title,
//         date length + date, number of sequences N, total residues
//         (64-bit, little-endian), longest sequence, then N+1 header offsets
//         (into .phr) and N+1 sequence offsets (into .psq)
//   .psq  ncbistdaa residues; sequence i is [seq_off[i], seq_off[i+1] - 1),
//         the byte before the next offset being a 0 separator
//   .pal  alias file; its "DBLIST v1 v2 ..." line names the volumes
//
// The header walk up to N follows the reference's own reader
// (functions.cpp:181-215: version, type == 1, title, date, N); the alias
// walk follows functions.cpp:217-353 (the first "DBLIST" line, names separated
// by spaces).  bpx_volume_write produces the same layout (with empty .phr
// headers), for tests and for packing synthetic databases.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bpx.h"

namespace {

thread_local std::string g_vol_err;

int vfail(int code, const std::string& msg) {
    g_vol_err = msg;
    return code;
}

bool read_file(const std::string& path, std::vector<uint8_t>* out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out->resize(n > 0 ? (size_t)n : 0);
    const size_t got = n > 0 ? std::fread(out->data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == out->size();
}

struct Cursor {
    const std::vector<uint8_t>& b;
    size_t at = 0;
    bool ok = true;
    uint32_t be32() {
        if (at + 4 > b.size()) {
            ok = false;
            return 0;
        }
        const uint32_t v = ((uint32_t)b[at] << 24) | ((uint32_t)b[at + 1] << 16) | ((uint32_t)b[at + 2] << 8) | b[at + 3];
        at += 4;
        return v;
    }
    uint64_t le64() {
        if (at + 8 > b.size()) {
            ok = false;
            return 0;
        }
        uint64_t v = 0;
        for (int i = 7; i >= 0; --i) v = (v << 8) | b[at + i];
        at += 8;
        return v;
    }
    std::string str(uint32_t n) {
        if (at + n > b.size()) {
            ok = false;
            return {};
        }
        std::string s(reinterpret_cast<const char*>(b.data() + at), n);
        at += n;
        return s;
    }
};

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

}  // namespace

struct bpx_volume {
    std::string title, date;
    uint32_t version = 0;
    uint32_t n = 0;
    uint64_t total = 0;
    uint32_t max_len = 0;
    std::vector<uint8_t> residues;  // concatenated, no separators
    std::vector<uint64_t> offsets;  // n + 1
    std::vector<uint32_t> lengths;  // n
};

extern "C" {

const char* bpx_volume_last_error(void) { return g_vol_err.c_str(); }

static int bpx_volume_read_impl(const char* base, bpx_volume** out) {
    if (!base || !out) return vfail(BPX_E_ARG, "NULL argument");
    *out = nullptr;
    std::vector<uint8_t> pin, psq;
    const std::string b(base);
    if (!read_file(b + ".pin", &pin)) return vfail(BPX_E_ARG, "cannot read " + b + ".pin");
    if (!read_file(b + ".psq", &psq)) return vfail(BPX_E_ARG, "cannot read " + b + ".psq");
    Cursor c{pin};
    auto* v = new (std::nothrow) bpx_volume;
    if (!v) return vfail(BPX_E_NOMEM, "host allocation failed");
    v->version = c.be32();
    const uint32_t type = c.be32();
    v->title = c.str(c.be32());
    v->date = c.str(c.be32());
    v->n = c.be32();
    v->total = c.le64();
    v->max_len = c.be32();
    if (!c.ok || v->version != 4 || type != 1) {
        const uint32_t version = v->version;
        delete v;
        return vfail(BPX_E_ARG, b + ".pin: not a version-4 protein index (version " + std::to_string(version) +
                                    ", type " + std::to_string(type) + ")");
    }
    // header offsets (unused here), then sequence offsets: two arrays of
    // n + 1 big-endian words, which the rest of the file must hold before
    // anything is sized by n (n = 2^32 - 1 would otherwise wrap n + 1)
    const uint64_t n1 = (uint64_t)v->n + 1;
    if (pin.size() < c.at || pin.size() - c.at < 8 * n1) {
        const uint32_t n = v->n;
        delete v;
        return vfail(BPX_E_ARG, b + ".pin: truncated offset arrays (" + std::to_string(n) + " sequences)");
    }
    c.at += 4 * n1;
    std::vector<uint32_t> so(n1);
    for (uint64_t i = 0; i < n1; ++i) so[i] = c.be32();
    v->offsets.assign(n1, 0);
    v->lengths.assign(v->n, 0);
    uint64_t total = 0;
    uint32_t mx = 0;
    for (uint32_t i = 0; i < v->n; ++i) {
        // sequence i is [so[i], so[i+1] - 1) followed by one separator byte;
        // compared in 64 bits so so[i] = 2^32 - 1 cannot wrap
        if ((uint64_t)so[i] + 1 > so[i + 1] || so[i + 1] > psq.size()) {
            delete v;
            return vfail(BPX_E_ARG, b + ".psq: sequence " + std::to_string(i) + " outside the file");
        }
        const uint32_t len = so[i + 1] - 1 - so[i];
        v->lengths[i] = len;
        v->offsets[i] = total;
        total += len;
        mx = std::max(mx, len);
    }
    v->offsets[v->n] = total;
    if (total != v->total || mx != v->max_len) {
        delete v;
        return vfail(BPX_E_ARG, b + ".pin: residue count / longest sequence disagree with the offsets");
    }
    v->residues.resize(total);
    for (uint32_t i = 0; i < v->n; ++i) {
        const uint8_t* src = psq.data() + so[i];
        for (uint32_t k = 0; k < v->lengths[i]; ++k)
            if (src[k] >= BPX_ALPHABET) {
                delete v;
                return vfail(BPX_E_ARG, b + ".psq: residue code >= 28 in sequence " + std::to_string(i));
            }
        std::memcpy(v->residues.data() + v->offsets[i], src, v->lengths[i]);
    }
    *out = v;
    return BPX_OK;
}

int bpx_volume_info(const bpx_volume* v, uint32_t* num_sequences, uint64_t* num_residues, uint32_t* max_length,
                    const char** title) {
    if (!v) return vfail(BPX_E_ARG, "volume is NULL");
    if (num_sequences) *num_sequences = v->n;
    if (num_residues) *num_residues = v->total;
    if (max_length) *max_length = v->max_len;
    if (title) *title = v->title.c_str();
    return BPX_OK;
}

int bpx_volume_arrays(const bpx_volume* v, const uint8_t** residues, const uint64_t** offsets,
                      const uint32_t** lengths) {
    if (!v) return vfail(BPX_E_ARG, "volume is NULL");
    if (residues) *residues = v->residues.data();
    if (offsets) *offsets = v->offsets.data();
    if (lengths) *lengths = v->lengths.data();
    return BPX_OK;
}

int bpx_volume_destroy(bpx_volume* v) {
    delete v;
    return BPX_OK;
}

static int bpx_volume_write_impl(const char* base, const char* title, const uint8_t* residues, const uint64_t* offsets,
                     const uint32_t* lengths, uint32_t n) {
    if (!base || (n && (!residues || !offsets || !lengths))) return vfail(BPX_E_ARG, "NULL argument");
    std::vector<uint8_t> psq{0}, pin, phr;
    std::vector<uint32_t> so(n + 1), ho(n + 1, 0);
    uint64_t total = 0;
    uint32_t mx = 0;
    for (uint32_t i = 0; i < n; ++i) {
        so[i] = (uint32_t)psq.size();
        psq.insert(psq.end(), residues + offsets[i], residues + offsets[i] + lengths[i]);
        psq.push_back(0);
        total += lengths[i];
        mx = std::max(mx, lengths[i]);
        if (psq.size() > 0xFFFFFFFFull) return vfail(BPX_E_ARG, "volume exceeds 4 GiB");
    }
    so[n] = (uint32_t)psq.size();
    const std::string t = title ? title : "";
    const std::string date = "Jan 1, 2025  0:00 AM";
    put_be32(pin, 4);
    put_be32(pin, 1);
    put_be32(pin, (uint32_t)t.size());
    pin.insert(pin.end(), t.begin(), t.end());
    put_be32(pin, (uint32_t)date.size());
    pin.insert(pin.end(), date.begin(), date.end());
    put_be32(pin, n);
    for (int i = 0; i < 8; ++i) pin.push_back((uint8_t)(total >> (8 * i)));
    put_be32(pin, mx);
    for (uint32_t x : ho) put_be32(pin, x);
    for (uint32_t x : so) put_be32(pin, x);
    const std::string b(base);
    for (auto& kv : {std::make_pair(std::string(".pin"), &pin), std::make_pair(std::string(".psq"), &psq),
                     std::make_pair(std::string(".phr"), &phr)}) {
        FILE* f = std::fopen((b + kv.first).c_str(), "wb");
        if (!f) return vfail(BPX_E_ARG, "cannot write " + b + kv.first);
        const size_t w = kv.second->empty() ? 0 : std::fwrite(kv.second->data(), 1, kv.second->size(), f);
        std::fclose(f);
        if (w != kv.second->size()) return vfail(BPX_E_ARG, "short write to " + b + kv.first);
    }
    return BPX_OK;
}

// Volume names of a database: the DBLIST line of base.pal (relative names
// are taken relative to the alias file's directory), else base itself.
// names receives them separated by '\n' (capacity bytes including the NUL);
// *needed gets the required capacity.
static int bpx_db_volumes_impl(const char* base, char* names, size_t capacity, size_t* needed, uint32_t* count) {
    if (!base) return vfail(BPX_E_ARG, "NULL argument");
    std::vector<std::string> vols;
    const std::string b(base);
    std::vector<uint8_t> pal;
    if (read_file(b + ".pal", &pal)) {
        const std::string dir = b.find('/') == std::string::npos ? "" : b.substr(0, b.rfind('/') + 1);
        std::string text(pal.begin(), pal.end());
        size_t at = 0;
        while (at < text.size()) {
            size_t eol = text.find('\n', at);
            if (eol == std::string::npos) eol = text.size();
            std::string line = text.substr(at, eol - at);
            at = eol + 1;
            if (line.compare(0, 6, "DBLIST") != 0) continue;
            size_t p = 6;
            while (p < line.size()) {
                while (p < line.size() && (line[p] == ' ' || line[p] == '\t' || line[p] == '\r')) ++p;
                size_t e = p;
                while (e < line.size() && line[e] != ' ' && line[e] != '\t' && line[e] != '\r') ++e;
                if (e > p) {
                    std::string name = line.substr(p, e - p);
                    if (name.size() > 1 && name.front() == '"' && name.back() == '"') name = name.substr(1, name.size() - 2);
                    vols.push_back(name[0] == '/' ? name : dir + name);
                }
                p = e;
            }
            break;
        }
        if (vols.empty()) return vfail(BPX_E_ARG, b + ".pal has no DBLIST line");
    } else {
        vols.push_back(b);
    }
    std::string joined;
    for (size_t i = 0; i < vols.size(); ++i) joined += (i ? "\n" : "") + vols[i];
    if (needed) *needed = joined.size() + 1;
    if (count) *count = (uint32_t)vols.size();
    if (names) {
        if (capacity < joined.size() + 1) return vfail(BPX_E_ARG, "names capacity too small");
        std::memcpy(names, joined.c_str(), joined.size() + 1);
    }
    return BPX_OK;
}

// Every volume of a database (alias or single volume), concatenated in
// DBLIST order, loaded as one resident arena: subject i of the result is
// the i-th sequence over the volumes (NCBI's OID order).
static int bpx_db_load_volumes_impl(bpx_ctx* ctx, const char* base, bpx_db** out) {
    if (!ctx || !base || !out) return vfail(BPX_E_ARG, "NULL argument");
    size_t need = 0;
    uint32_t nv = 0;
    int rc = bpx_db_volumes(base, nullptr, 0, &need, &nv);
    if (rc) return rc;
    std::string names(need, '\0');
    rc = bpx_db_volumes(base, &names[0], need, nullptr, nullptr);
    if (rc) return rc;
    std::vector<uint8_t> res;
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
    size_t at = 0;
    while (at < names.size() && names[at]) {
        size_t e = names.find('\n', at);
        if (e == std::string::npos) e = std::strlen(names.c_str());
        const std::string vol = names.substr(at, e - at);
        at = e + 1;
        bpx_volume* v = nullptr;
        rc = bpx_volume_read(vol.c_str(), &v);
        if (rc) return rc;
        const uint64_t base_off = res.size();
        res.insert(res.end(), v->residues.begin(), v->residues.end());
        for (uint32_t i = 0; i < v->n; ++i) {
            off.push_back(base_off + v->offsets[i]);
            len.push_back(v->lengths[i]);
        }
        bpx_volume_destroy(v);
        if (e >= names.size() || names[e] == '\0') break;
    }
    if (len.size() > 0xFFFFFFFFull) return vfail(BPX_E_ARG, "more than 2^32 sequences");
    return bpx_db_load(ctx, res.data(), off.data(), len.data(), (uint32_t)len.size(), out);
}

// ---- the reference's own GPU database files (.gpu / .gpuinfo) -------------
//
// .gpuinfo is text (Write_GPU_information_file_preliminary / _file,
// functions.cpp:1-46): a comment line before every value; blocks, threads,
// percentage, number of volumes; then per volume its .gpu file name, byte
// size, sequences in the volume, sequences in the .gpu file, store limit,
// break limit, number of groups and each group's length.  A .gpu file
// (CreateGPU_BLASTP_database, functions.cpp:364-613) holds the first
// sequences of the volume (ascending length) in groups of G = blocks *
// threads: group g spans len_g * G bytes (len_g: its longest sequence rounded
// up to 64), and residue 4i + j of the group's m-th sequence sits at byte
// (m + i * G) * 4 + j of the group; the rest is zero.  The files hold no
// per-sequence length: protein residues are never 0, so a sequence ends at
// its last non-zero byte.

namespace {

struct GpuInfo {
    int64_t blocks = 0, threads = 0, percentage = 0;
    struct Vol {
        std::string name;
        uint64_t bytes = 0;
        int64_t volume_length = 0, num_sequences = 0, store_limit = 0, break_limit = 0;
        std::vector<int64_t> group_len;
    };
    std::vector<Vol> vols;
};

int parse_gpuinfo(const std::string& path, GpuInfo* gi) {
    std::vector<uint8_t> raw;
    if (!read_file(path, &raw)) return vfail(BPX_E_ARG, "cannot read " + path);
    std::vector<std::string> vals;  // every non-comment line, in order
    std::string text(raw.begin(), raw.end());
    size_t at = 0;
    while (at < text.size()) {
        size_t eol = text.find('\n', at);
        if (eol == std::string::npos) eol = text.size();
        std::string line = text.substr(at, eol - at);
        at = eol + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (!line.empty() && line[0] == '#') continue;
        vals.push_back(line);
    }
    size_t k = 0;
    bool ok = true;
    auto num = [&](int64_t* v) {
        if (k >= vals.size()) {
            ok = false;
            return;
        }
        char* end = nullptr;
        const long long x = std::strtoll(vals[k].c_str(), &end, 10);
        if (end == vals[k].c_str()) ok = false;
        *v = x;
        ++k;
    };
    int64_t nv = 0;
    num(&gi->blocks);
    num(&gi->threads);
    num(&gi->percentage);
    num(&nv);
    if (!ok || gi->blocks <= 0 || gi->threads <= 0 || nv < 0 || nv > 1000000)
        return vfail(BPX_E_ARG, path + ": malformed header (blocks, threads, percentage, volumes)");
    for (int64_t v = 0; v < nv && ok; ++v) {
        GpuInfo::Vol vol;
        if (k >= vals.size()) {
            ok = false;
            break;
        }
        vol.name = vals[k++];
        int64_t bytes = 0, groups = 0;
        num(&bytes);
        num(&vol.volume_length);
        num(&vol.num_sequences);
        num(&vol.store_limit);
        num(&vol.break_limit);
        num(&groups);
        if (!ok || bytes < 0 || groups < 0 || groups > (1 << 30) || vol.num_sequences < 0) {
            ok = false;
            break;
        }
        vol.bytes = (uint64_t)bytes;
        vol.group_len.resize((size_t)groups);
        for (auto& g : vol.group_len) num(&g);
        gi->vols.push_back(std::move(vol));
    }
    if (!ok) return vfail(BPX_E_ARG, path + ": malformed volume record");
    return BPX_OK;
}

}  // namespace

static int bpx_gpu_info_impl(const char* gpuinfo, uint32_t* num_volumes, uint32_t* group_size, uint32_t* percentage) {
    if (!gpuinfo) return vfail(BPX_E_ARG, "NULL argument");
    GpuInfo gi;
    const int rc = parse_gpuinfo(gpuinfo, &gi);
    if (rc) return rc;
    if (num_volumes) *num_volumes = (uint32_t)gi.vols.size();
    if (group_size) *group_size = (uint32_t)(gi.blocks * gi.threads);
    if (percentage) *percentage = (uint32_t)gi.percentage;
    return BPX_OK;
}

static int bpx_gpu_volume_read_impl(const char* gpuinfo, uint32_t volume, bpx_volume** out) {
    if (!gpuinfo || !out) return vfail(BPX_E_ARG, "NULL argument");
    *out = nullptr;
    GpuInfo gi;
    int rc = parse_gpuinfo(gpuinfo, &gi);
    if (rc) return rc;
    if (volume >= gi.vols.size()) return vfail(BPX_E_ARG, "volume index out of range");
    const GpuInfo::Vol& vol = gi.vols[volume];
    const uint64_t G = (uint64_t)(gi.blocks * gi.threads);
    const uint64_t n = (uint64_t)vol.num_sequences;
    const uint64_t ngroups = vol.group_len.size();
    if (ngroups != (n + G - 1) / G) return vfail(BPX_E_ARG, "group count does not match the sequence count");
    uint64_t bytes = 0;
    for (int64_t L : vol.group_len) {
        if (L < 0 || L > (1 << 30)) return vfail(BPX_E_ARG, "bad group length");
        bytes += (uint64_t)L * G;
    }
    if (bytes != vol.bytes) return vfail(BPX_E_ARG, "volume size does not match its group lengths");
    // the name as written; else the same file name next to the .gpuinfo
    std::vector<uint8_t> db;
    std::string name = vol.name;
    if (!read_file(name, &db)) {
        const std::string info(gpuinfo);
        const std::string dir = info.find('/') == std::string::npos ? "" : info.substr(0, info.rfind('/') + 1);
        const std::string leaf = name.find('/') == std::string::npos ? name : name.substr(name.rfind('/') + 1);
        name = dir + leaf;
        if (!read_file(name, &db)) return vfail(BPX_E_ARG, "cannot read " + vol.name);
    }
    if (db.size() != bytes) return vfail(BPX_E_ARG, name + ": size differs from the .gpuinfo record");
    auto* v = new (std::nothrow) bpx_volume;
    if (!v) return vfail(BPX_E_NOMEM, "out of memory");
    v->version = 0;
    v->n = (uint32_t)n;
    v->offsets.reserve(n + 1);
    v->lengths.reserve(n);
    uint64_t gbase = 0;
    for (uint64_t g = 0; g < ngroups; ++g) {
        const uint64_t L = (uint64_t)vol.group_len[g];
        const uint64_t members = std::min<uint64_t>(G, n - g * G);
        for (uint64_t m = 0; m < members; ++m) {
            auto at = [&](uint64_t r) { return gbase + (m + (r / 4) * G) * 4 + (r & 3); };
            uint64_t len = L;
            while (len > 0 && db[at(len - 1)] == 0) --len;
            v->offsets.push_back(v->residues.size());
            v->lengths.push_back((uint32_t)len);
            for (uint64_t r = 0; r < len; ++r) v->residues.push_back(db[at(r)]);
            v->max_len = std::max<uint32_t>(v->max_len, (uint32_t)len);
        }
        gbase += L * G;
    }
    v->offsets.push_back(v->residues.size());
    v->total = v->residues.size();
    *out = v;
    return BPX_OK;
}


// The readers size host arrays from file contents: an allocation failure
// is reported as BPX_E_NOMEM, never thrown across the C ABI.
int bpx_volume_read(const char* base, bpx_volume** out) {
    try {
        return bpx_volume_read_impl(base, out);
    } catch (const std::bad_alloc&) {
        return vfail(BPX_E_NOMEM, "host allocation failed");
    }
}

int bpx_volume_write(const char* base, const char* title, const uint8_t* residues, const uint64_t* offsets,
                     const uint32_t* lengths, uint32_t n) {
    try {
        return bpx_volume_write_impl(base, title, residues, offsets, lengths, n);
    } catch (const std::bad_alloc&) {
        return vfail(BPX_E_NOMEM, "host allocation failed");
    }
}

int bpx_db_volumes(const char* base, char* names, size_t capacity, size_t* needed, uint32_t* count) {
    try {
        return bpx_db_volumes_impl(base, names, capacity, needed, count);
    } catch (const std::bad_alloc&) {
        return vfail(BPX_E_NOMEM, "host allocation failed");
    }
}

int bpx_db_load_volumes(bpx_ctx* ctx, const char* base, bpx_db** out) {
    try {
        return bpx_db_load_volumes_impl(ctx, base, out);
    } catch (const std::bad_alloc&) {
        return vfail(BPX_E_NOMEM, "host allocation failed");
    }
}

int bpx_gpu_info(const char* gpuinfo, uint32_t* num_volumes, uint32_t* group_size, uint32_t* percentage) {
    try {
        return bpx_gpu_info_impl(gpuinfo, num_volumes, group_size, percentage);
    } catch (const std::bad_alloc&) {
        return vfail(BPX_E_NOMEM, "host allocation failed");
    }
}

int bpx_gpu_volume_read(const char* gpuinfo, uint32_t volume, bpx_volume** out) {
    try {
        return bpx_gpu_volume_read_impl(gpuinfo, volume, out);
    } catch (const std::bad_alloc&) {
        return vfail(BPX_E_NOMEM, "host allocation failed");
    }
}

}  // extern "C"
