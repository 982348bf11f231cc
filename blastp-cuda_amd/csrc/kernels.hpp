// kernels.hpp -- launch interface between the C-ABI runtime and the gfx950
// word-finder kernels.  Host-visible structs only; no HIP types leak into
// the public header.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "bpx.h"

namespace bpx {

// Word-finder launch for one query over a list of subjects.
struct ScanParams {
    // database arena
    const uint8_t* residues;
    const uint64_t* soff;       // per subject: arena offset
    const uint32_t* slen;       // per subject: length
    const uint32_t* order;      // work list of absolute subject ids
    uint32_t norder;
    uint32_t first;             // subject id of stats/rec_loc slot 0
    // query table (packed: groups | starts | entries)
    const uint32_t* table;
    uint32_t table_words;
    uint32_t ngroups;
    uint32_t npresent;
    const uint8_t* query;
    int32_t qlen;
    const int8_t* matrix;       // 28 x 28, row = query residue
    int32_t word, window, xdrop, cutoff, diag_mask;
    // scheduling
    uint32_t* work_counter;
    uint32_t grab;
    // outputs
    int32_t* stats;             // 3 per subject slot
    uint32_t* rec_loc;          // per subject slot: first record in seg
    bpx_init_hsp* seg;          // per-wave record segments
    uint32_t seg_cap;           // records per wave segment
    uint32_t* overflow;         // bit 0: a segment filled; bit 1: queue too small
    // LDS layout (bytes)
    uint32_t lds_table, lds_matrix, lds_query, lds_wave, wave_stride, qcap;
    uint32_t lds_total;
};

// LDS bytes the word finder needs per workgroup; table_in_lds selects
// whether the query table is staged.
struct LdsPlan {
    uint32_t table, matrix, query, wave, wave_stride, total;
    bool table_in_lds;
};
LdsPlan plan_lds(uint32_t table_bytes, int32_t qlen, int32_t diag_len, uint32_t qcap,
                 int waves, uint32_t lds_limit, bool allow_table_in_lds);

hipError_t launch_scan(const ScanParams& p, bool table_in_lds, int waves, int groups,
                       hipStream_t stream);

// Compaction: per-subject saved counts -> CSR offsets -> ordered records.
struct CompactParams {
    const int32_t* stats;       // 3 per subject slot
    const uint32_t* rec_loc;
    const bpx_init_hsp* seg;
    uint32_t count;             // subject slots
    uint32_t first;             // absolute id of slot 0
    uint32_t* tile_sum;         // ceil(count / kTile)
    uint32_t* rec_off;          // count + 1
    bpx_init_hsp* records;      // capacity >= total
    uint32_t* subjects;         // capacity >= total
    unsigned long long* ustats; // 5 sums (bpx_ungapped_stats order) + total records
    const uint32_t* overflow;
};
constexpr uint32_t kTile = 1024;
hipError_t launch_compact(const CompactParams& p, hipStream_t stream);

hipError_t launch_fill_u32(uint32_t* dst, uint32_t value, uint64_t n, hipStream_t stream);

}  // namespace bpx
