// table.hpp -- host-side query lookup table for the gfx950 word finder.
//
// The reference consumes NCBI's BlastAaLookupTable in its "smallbone" form:
// a presence vector plus 12-byte cells of up to 3 inline u16 query offsets
// and an overflow array (gpu_blastp.cu:45-61, read at
// gpu_blastp_kernel.cu:182-192).  That backbone is 28,540 cells x 12 B and
// does not fit in LDS, so this table keeps the same information in a
// rank-indexed CSR that does:
//
//   groups[g]   = {pv bits of words 32g..32g+31, number of present words
//                  in groups < g}                       (8 B per 32 words)
//   starts[r]   = first entry of the r-th present word  (u32, npresent+1)
//   entries[e]  = query offset                          (u16)
//
// A word w is present iff bit (w & 31) of groups[w>>5].pv is set; its rank
// is groups[w>>5].base + popcount(pv & ((1<<(w&31))-1)); its query offsets
// are entries[starts[rank] .. starts[rank+1]).
#pragma once

#include <cstdint>
#include <vector>

namespace bpx {

constexpr int kAlphabet = 28;
constexpr int kCharsize = 5;

// Presence words a packed table keeps for words of word_length residues:
// those covering codes whose first residue is < 28 (of the 5-bit codes).
inline uint32_t packed_presence_groups(int32_t word_length) {
    return ((28u << (5 * (word_length - 1))) + 31u) / 32u;
}

struct HostTable {
    int32_t word_length = 0;
    uint32_t nwords = 0;                // 1 << (5 * word_length)
    uint32_t ngroups = 0;               // nwords / 32
    std::vector<uint32_t> groups;       // 2 * ngroups: {pv, base} pairs
    std::vector<uint32_t> starts;       // npresent + 1
    std::vector<uint16_t> entries;      // query offsets, ascending per word
    uint32_t npresent = 0;
    uint32_t max_cell = 0;

    // Serialises the table into one word-aligned blob:
    // [groups (8*ngroups B)][starts][entries (2*n B, padded to 4)], where the
    // starts are u16 (padded to 4 B) when the table has at most 65,535
    // entries and u32 otherwise.
    bool narrow_starts() const;
    // Presence words kept in the packed table: only words whose first
    // residue code is < 28 can occur, so the rest (always empty) are dropped.
    uint32_t packed_groups() const;
    std::vector<uint32_t> pack() const;
    uint64_t packed_bytes() const;
    // Start offsets indexed by word over the packed word range W
    // (packed_groups() * 32), W + 2 entries: word w's query offsets are
    // entries[d[w] .. d[w+1]), empty for an absent word, and d[W] = d[W+1] =
    // the entry count, so any word at or past W reads an empty range.  A
    // table read from HBM is looked up through this array in one step
    // instead of presence word, then rank, then starts.
    std::vector<uint32_t> direct_starts() const;
};

// Builds the neighbourhood table of `query` (ncbistdaa, qlen >= 0): word w is
// listed under query position p iff sum_i M[q[p+i]][w_i] >= threshold, or w
// is the query word itself.  Neighbours are enumerated with row-maximum
// pruning over the 28-letter alphabet, once per distinct query word.
// Returns false on bad arguments.
bool build_table(const uint8_t* query, int32_t qlen, const int8_t* matrix28,
                 int32_t threshold, int32_t word_length, HostTable* out);

// One table for several queries scanned together: query k occupies
// positions [base_k, base_k + len_k) of their concatenation and its words
// are listed under base_k + p.  Words never span two queries.  The
// concatenation must be shorter than 65,536 residues.
struct QuerySpan {
    const uint8_t* seq;
    int32_t len;
};
bool build_table_multi(const QuerySpan* queries, int32_t nq, const int8_t* matrix28,
                       int32_t threshold, int32_t word_length, HostTable* out);

}  // namespace bpx
