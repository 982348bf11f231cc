// table.cpp -- neighbourhood enumeration and CSR packing of the query table.
#include "table.hpp"

#include "bpx.h"

#include <algorithm>
#include <cstring>
#include <unordered_map>

namespace bpx {

namespace {

inline uint32_t word_index(const uint8_t* w, int word_length) {
    uint32_t idx = 0;
    for (int i = 0; i < word_length; ++i) idx = (idx << kCharsize) | w[i];
    return idx;
}

// Calls emit(idx) for every word whose score against `qw` is >= threshold.
// Pruned: a prefix is abandoned as soon as even the best remaining letters
// (row maxima) cannot lift it to the threshold.
template <class Emit>
void enumerate_neighbours(const uint8_t* qw, int word_length, const int8_t* M,
                          const int32_t* row_max, int32_t threshold, Emit emit) {
    const int8_t* r0 = M + qw[0] * kAlphabet;
    const int8_t* r1 = M + qw[1] * kAlphabet;
    if (word_length == 2) {
        for (int a = 0; a < kAlphabet; ++a) {
            const int32_t s1 = r0[a];
            if (s1 + row_max[qw[1]] < threshold) continue;
            for (int b = 0; b < kAlphabet; ++b)
                if (s1 + r1[b] >= threshold) emit((uint32_t)((a << kCharsize) | b));
        }
        return;
    }
    const int8_t* r2 = M + qw[2] * kAlphabet;
    const int32_t tail12 = row_max[qw[1]] + row_max[qw[2]];
    const int32_t tail2 = row_max[qw[2]];
    for (int a = 0; a < kAlphabet; ++a) {
        const int32_t s1 = r0[a];
        if (s1 + tail12 < threshold) continue;
        for (int b = 0; b < kAlphabet; ++b) {
            const int32_t s2 = s1 + r1[b];
            if (s2 + tail2 < threshold) continue;
            for (int c = 0; c < kAlphabet; ++c)
                if (s2 + r2[c] >= threshold)
                    emit((uint32_t)((a << (2 * kCharsize)) | (b << kCharsize) | c));
        }
    }
}

}  // namespace

bool build_table(const uint8_t* query, int32_t qlen, const int8_t* M,
                 int32_t threshold, int32_t word_length, HostTable* out) {
    QuerySpan one{query, qlen};
    return build_table_multi(&one, 1, M, threshold, word_length, out);
}

bool build_table_multi(const QuerySpan* queries, int32_t nq, const int8_t* M,
                       int32_t threshold, int32_t word_length, HostTable* out) {
    if (!out || !M || word_length < 2 || word_length > 3 || nq < 0 || (nq > 0 && !queries)) return false;
    int64_t total = 0;
    for (int32_t k = 0; k < nq; ++k) {
        const QuerySpan& q = queries[k];
        if (q.len < 0 || (q.len > 0 && !q.seq)) return false;
        for (int32_t i = 0; i < q.len; ++i)
            if (q.seq[i] >= kAlphabet) return false;
        total += q.len;
    }
    if (total > 65535) return false;

    HostTable t;
    t.word_length = word_length;
    t.nwords = 1u << (kCharsize * word_length);
    t.ngroups = t.nwords / 32;

    int32_t row_max[kAlphabet];
    for (int i = 0; i < kAlphabet; ++i) {
        int32_t m = M[i * kAlphabet];
        for (int j = 1; j < kAlphabet; ++j) m = std::max<int32_t>(m, M[i * kAlphabet + j]);
        row_max[i] = m;
    }

    // Distinct query words -> ascending positions in the concatenation.
    std::unordered_map<uint32_t, std::vector<uint16_t>> by_word;
    std::vector<uint32_t> word_order;  // first-occurrence order, for determinism
    int32_t base = 0;
    for (int32_t k = 0; k < nq; ++k) {
        const QuerySpan& q = queries[k];
        const int32_t npos = q.len - word_length + 1;
        for (int32_t p = 0; p < npos; ++p) {
            const uint32_t w = word_index(q.seq + p, word_length);
            auto it = by_word.find(w);
            if (it == by_word.end()) {
                word_order.push_back(w);
                by_word.emplace(w, std::vector<uint16_t>{(uint16_t)(base + p)});
            } else {
                it->second.push_back((uint16_t)(base + p));
            }
        }
        base += q.len;
    }

    // Count, then fill, per neighbour word.
    std::vector<uint32_t> count(t.nwords + 1, 0);
    std::vector<uint32_t> neigh;
    std::vector<std::vector<uint32_t>> neigh_of(word_order.size());
    for (size_t k = 0; k < word_order.size(); ++k) {
        const uint32_t w = word_order[k];
        uint8_t qw[3];
        for (int i = 0; i < word_length; ++i)
            qw[i] = (uint8_t)((w >> (kCharsize * (word_length - 1 - i))) & 31u);
        neigh.clear();
        bool self_found = false;
        enumerate_neighbours(qw, word_length, M, row_max, threshold, [&](uint32_t idx) {
            neigh.push_back(idx);
            self_found |= (idx == w);
        });
        if (!self_found) neigh.push_back(w);  // the exact word is always indexed
        const uint32_t npos_w = (uint32_t)by_word[w].size();
        for (uint32_t idx : neigh) count[idx] += npos_w;
        neigh_of[k] = neigh;
    }

    t.groups.assign(2 * (size_t)t.ngroups, 0);
    uint32_t present = 0;
    for (uint32_t g = 0; g < t.ngroups; ++g) {
        uint32_t bits = 0;
        for (uint32_t j = 0; j < 32; ++j)
            if (count[g * 32 + j]) bits |= 1u << j;
        t.groups[2 * g] = bits;
        t.groups[2 * g + 1] = present;
        present += (uint32_t)__builtin_popcount(bits);
    }
    t.npresent = present;
    t.starts.assign((size_t)present + 1, 0);
    std::vector<uint32_t> rank_of(t.nwords, UINT32_MAX);
    {
        uint32_t r = 0, acc = 0;
        for (uint32_t w = 0; w < t.nwords; ++w)
            if (count[w]) {
                rank_of[w] = r;
                t.starts[r++] = acc;
                acc += count[w];
                t.max_cell = std::max(t.max_cell, count[w]);
            }
        t.starts[present] = acc;
        t.entries.assign(acc, 0);
    }
    std::vector<uint32_t> fill(present, 0);
    for (size_t k = 0; k < word_order.size(); ++k) {
        const auto& pos = by_word[word_order[k]];
        for (uint32_t idx : neigh_of[k]) {
            const uint32_t r = rank_of[idx];
            uint16_t* dst = t.entries.data() + t.starts[r] + fill[r];
            std::memcpy(dst, pos.data(), pos.size() * sizeof(uint16_t));
            fill[r] += (uint32_t)pos.size();
        }
    }
    // Ascending query offsets within each word: the scan order of the
    // sequential definition (positions outer, words inner).
    for (uint32_t r = 0; r < present; ++r)
        std::sort(t.entries.begin() + t.starts[r], t.entries.begin() + t.starts[r + 1]);

    *out = std::move(t);
    return true;
}

bool HostTable::narrow_starts() const { return entries.size() <= 0xFFFFu; }

uint32_t HostTable::packed_groups() const { return packed_presence_groups(word_length); }

uint64_t HostTable::packed_bytes() const {
    const uint64_t e = ((uint64_t)entries.size() * 2 + 3) & ~3ull;
    const uint64_t st = narrow_starts() ? (2ull * (npresent + 1) + 3) & ~3ull : 4ull * (npresent + 1);
    return 8ull * packed_groups() + st + e;
}

std::vector<uint32_t> HostTable::direct_starts() const {
    const uint32_t W = packed_groups() * 32u;
    std::vector<uint32_t> d((size_t)W + 2);
    uint32_t r = 0;
    for (uint32_t w = 0; w < W; ++w) {
        d[w] = starts[r];
        if ((groups[2 * (w >> 5)] >> (w & 31)) & 1u) ++r;
    }
    d[W] = d[(size_t)W + 1] = starts[npresent];
    return d;
}

std::vector<uint32_t> HostTable::pack() const {
    std::vector<uint32_t> blob(packed_bytes() / 4, 0);
    uint8_t* b = reinterpret_cast<uint8_t*>(blob.data());
    std::memcpy(b, groups.data(), 8ull * packed_groups());
    b += 8ull * packed_groups();
    if (narrow_starts()) {
        uint16_t* s16 = reinterpret_cast<uint16_t*>(b);
        for (uint32_t r = 0; r <= npresent; ++r) s16[r] = (uint16_t)starts[r];
        b += (2ull * (npresent + 1) + 3) & ~3ull;
    } else {
        std::memcpy(b, starts.data(), 4ull * (npresent + 1));
        b += 4ull * (npresent + 1);
    }
    if (!entries.empty()) std::memcpy(b, entries.data(), 2ull * entries.size());
    return blob;
}

}  // namespace bpx

namespace bpx {
const int8_t* blosum62();
}

extern "C" int bpx_table_build(const uint8_t* query, int32_t qlen, const int8_t* matrix, int32_t threshold,
                               int32_t word_length, uint32_t* cell_start, uint16_t* entries, uint64_t capacity,
                               uint64_t* num_entries) {
    bpx::HostTable t;
    if (!bpx::build_table(query, qlen, matrix ? matrix : bpx::blosum62(), threshold, word_length, &t))
        return BPX_E_ARG;
    if (num_entries) *num_entries = t.entries.size();
    if (!cell_start) return BPX_E_ARG;
    if (capacity < t.entries.size()) return BPX_E_ARG;
    // the start offsets by word the scan reads for tables in HBM, extended
    // over the words it drops (first residue code >= 28: never present)
    const std::vector<uint32_t> d = t.direct_starts();
    const uint32_t W = (uint32_t)d.size() - 2;
    for (uint32_t w = 0; w <= t.nwords; ++w) cell_start[w] = d[std::min(w, W)];
    if (entries && !t.entries.empty()) std::memcpy(entries, t.entries.data(), t.entries.size() * 2);
    return BPX_OK;
}
