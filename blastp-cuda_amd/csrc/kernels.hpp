// kernels.hpp -- launch interface between the C-ABI runtime and the gfx950
// word-finder kernels.  Host-visible structs only.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "bpx.h"

namespace bpx {

// One query of a scan group (device array, 32 B).  A group is a set of
// queries scanned together against each subject through one combined lookup
// table; a single-query scan is a group of one.
struct QueryInfo {
    uint32_t qstart;   // offset of the query in its group's concatenated residues
    uint32_t qlen;
    uint32_t dbase;    // first slot of the query's diagonal table
    uint32_t dmask;    // diag_len - 1
    int32_t xdrop;
    int32_t cutoff;
    uint32_t qid;      // index of the query in the scan (row of the count arrays)
    uint32_t pad;
};

struct GroupInfo {
    uint32_t table_word;   // offset (in u32 words) of the group's packed table
    uint32_t table_words;
    uint32_t npresent;     // present words (rank space)
    uint32_t qres_off;     // byte offset of the group's concatenated residues
    uint32_t qres_len;
    uint32_t qinfo_first;  // first QueryInfo of the group
    uint32_t nq;           // queries in the group
    uint32_t diag_slots;   // sum of the members' diagonal-table lengths
    uint32_t flags;        // kGroupNarrowStarts: u16 start offsets
    uint32_t direct_word;  // offset (in u32 words, same arena) of its start offsets by word
    uint32_t pad[2];
};
constexpr uint32_t kGroupNarrowStarts = 1;

// A unit of dynamic work: one group against a run of the subject order.
struct WorkItem {
    uint32_t group;
    uint32_t first;   // index into the order list
    uint32_t count;
    uint32_t chunk;   // flat launches: subjects a wavefront claims at a time (0: the launch's)
};

// Record as staged in a wavefront's output segment: the HSP, the index of
// its query within the group and its subject's slot (subject id - first).
struct SegRecord {
    bpx_init_hsp h;
    uint32_t j;
    uint32_t slot;
};

struct ScanParams {
    // database arena
    const uint8_t* residues;
    const uint64_t* soff;
    const uint32_t* slen;
    const uint32_t* order;      // subject ids, processing order
    const uint64_t* poff;       // arena offset of order[k]
    const uint32_t* plen;       // length of order[k]
    uint32_t first;             // subject id of slot 0
    uint32_t count;             // subject slots
    // queries
    const GroupInfo* groups;
    const QueryInfo* qinfo;
    const uint32_t* tables;
    const uint8_t* qres;
    const uint8_t* qmap;        // per group position: index of its query in the group (beside qres)
    const int8_t* matrix;       // 28 x 28, row = query residue
    uint32_t ngroups_pv;        // presence-vector words (nwords / 32), same for all groups
    int32_t word, window;
    // work
    const WorkItem* items;
    uint32_t nitems;
    uint32_t* work_counter;
    // nonzero: every workgroup works through all items in order (one per
    // group), each wavefront claiming `flat` subjects at a time from the
    // item's counter in flat_counters; zero: workgroups claim items from
    // *work_counter and their waves claim the item's subjects one at a time
    uint32_t flat;
    uint32_t* flat_counters;
    // outputs
    uint32_t* cnt;              // saved records per (qid, slot): qid * count + slot; zero-filled
    int32_t* stats;             // optional: 3 per (qid, slot): qid * count + slot, in pinned host memory
    uint32_t* run_loc;          // per (group, slot): first segment record of the run
    uint32_t* run_len;          // per (group, slot): records in the run; zero-filled
    unsigned long long* ustats; // 5 per qid (bpx_ungapped_stats order)
    SegRecord* seg;
    uint32_t seg_cap;           // records per wave segment
    uint32_t seg_wave_base;     // first wave segment of this launch
    uint32_t* overflow;         // bit 0: a segment filled; bit 1: ring too small
    unsigned long long* phase;  // optional: 8 words per wave (launch-wide wave index, as the segments):
                                // start, end, staging, group changes, subjects and bytes staged
    uint32_t* max_fill;         // largest per-wave record count, when above seg_cap
    // LDS layout (bytes) and shape
    uint32_t lds_table, lds_qres, lds_qmap, lds_qinfo, lds_matrix, lds_wave, wave_stride;
    uint32_t lds_total;
    uint32_t table_cap_words;   // LDS table region (words); 0 = tables read from HBM
    uint32_t qres_cap;          // LDS bytes for concatenated residues / query map
    uint32_t max_nq;            // LDS slots for QueryInfo / per-query counters
    uint32_t diag_cap;          // diagonal slots per wave (u16 slots in narrow launches)
    uint32_t qcap;              // hit ring entries per wave (power of two)
    uint32_t claim_slots;       // hashed claim table entries per wave (power of two, >= 64)
    uint32_t sbuf;              // subject buffer bytes per wave (power of two)
};

struct LdsPlan {
    uint32_t table, qres, qmap, qinfo, matrix, wave, wave_stride, total;
    uint32_t table_cap_words, qres_cap, max_nq;
    uint32_t qcap, claim_slots;
};

constexpr uint32_t kTaskCap = 128;     // queued extensions per wave
constexpr uint32_t kDefCap = 128;      // deferred hits per wave (hits on pending slots)
// Deferred hits are re-applied right after a flush; each can queue at most
// one extension, and a flush keeps at most kTaskCap minus that many tasks
// queued, so they always fit.
static_assert(kDefCap <= kTaskCap, "re-applied deferred hits must fit the emptied extension queue");

// Narrow launches keep 16-bit diagonal slots and clear the table for every
// subject; they take subjects with length + window < kNarrowLimit.  Wide
// launches keep 32-bit slots with NCBI's running diagonal offset.
constexpr uint32_t kNarrowLimit = 1u << 14;
constexpr int kMaxWaves = 16;

LdsPlan plan_lds(uint32_t max_table_bytes, uint32_t max_qres, uint32_t max_nq, uint32_t diag_cap,
                 uint32_t qcap, uint32_t sbuf, uint32_t claim_slots, int waves, bool table_in_lds, bool narrow);

// Sets the kernel's dynamic-LDS limit for p.lds_total (done by launch_scan
// unless set_attr is false, e.g. while the stream is being captured).
hipError_t prepare_scan(const ScanParams& p, bool table_in_lds, bool multi, bool narrow);
hipError_t launch_scan(const ScanParams& p, bool table_in_lds, bool multi, bool narrow, int waves, int groups,
                       hipStream_t stream, bool set_attr = true);

// Compaction: counts -> offsets (in place) -> records in (query, subject,
// subject_offset, query_offset) order.
struct CompactParams {
    uint32_t* cnt;              // nq * count saved counts; overwritten by offsets
    uint64_t n;                 // nq * count
    uint32_t* tile_sum;
    const uint32_t* run_loc;
    const uint32_t* run_len;
    const GroupInfo* groups;
    const QueryInfo* qinfo;
    uint32_t ngroups;
    uint32_t count;             // subject slots
    uint32_t first;
    SegRecord* seg;
    bpx_init_hsp* records;
    uint32_t* subjects;
    uint32_t* queries;          // optional
    uint64_t rec_cap;           // entries of records / subjects / queries; records past it are counted, not written
    unsigned long long* total;  // total records
    const uint32_t* overflow;
};
constexpr uint32_t kTile = 1024;
hipError_t launch_compact(const CompactParams& p, hipStream_t stream);

constexpr int kZeroRanges = 4;
struct ZeroRanges {
    uint32_t* dst[kZeroRanges];
    uint64_t n[kZeroRanges];
    int count;
};
// zero-fills the ranges with one kernel launch (no memset nodes in a graph)
hipError_t launch_zero_u32(const ZeroRanges& z, hipStream_t stream);

}  // namespace bpx
