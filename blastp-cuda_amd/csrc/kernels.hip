// kernels.hip -- gfx950 kernels of the blastp word finder.
//
// k_scan: two-hit word finding + ungapped X-drop extension for one query.
// The reference (gpu_blastp_kernel.cu:42-320) gives each CUDA thread a whole
// subject and a private diagonal array in global memory, so every word hit is
// a random global read-modify-write.  Here a 64-lane wavefront owns a
// subject; its diagonal table lives in LDS, the query table is staged in LDS
// once per workgroup, and the lanes work on 64 subject positions at a time:
//
//   1. each lane forms the word at its position, tests the presence bits and
//      looks up its cell (rank into the packed CSR table);
//   2. a wave prefix sum over the cell sizes compacts the hits, in
//      (subject offset, query offset) order, into a per-wave LDS ring;
//   3. the ring is drained 64 hits at a time: each lane takes one hit and
//      claims its diagonal with an LDS atomic-min on its lane number, so hits
//      that share a diagonal are applied in scan order while hits on
//      different diagonals proceed together -- the exact sequential
//      semantics of the two-hit state machine (gpu_blastp_kernel.cu:194-301);
//   4. a lane whose hit completes a pair runs the X-drop extension
//      (s_BlastAaExtendTwoHit/Left/Right, gpu_blastp_kernel.cu:341-549) and
//      appends a record to its wave's private output segment.
//
// Records are later moved into subject order by k_tile_sums / k_tile_scan /
// k_gather (count, scan, copy: no atomics on the output).
#include "kernels.hpp"

#include <hip/hip_runtime.h>

namespace bpx {

namespace {

constexpr uint32_t kFlagBit = 0x80000000u;
constexpr int kMatrixStride = 32;  // LDS matrix rows padded 28 -> 32

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t rank_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

struct TwoHit {
    int32_t hsp_q, hsp_s, len, score, s_last;
    bool right_extend;
};

// s_BlastAaExtendTwoHit semantics: best prefix of the second word, left
// X-drop from one past it, right X-drop only if the left reached the first
// hit.  Subject bytes come from HBM (L1/L2-resident: the wave is streaming
// this subject), query bytes and scores from LDS.
__device__ __forceinline__ TwoHit extend_two_hit(const int8_t* __restrict__ M,
                                                 const uint8_t* __restrict__ Q, int32_t qlen,
                                                 const uint8_t* __restrict__ S, int32_t slen,
                                                 int32_t s_left, int32_t s, int32_t q,
                                                 int32_t xdrop, int32_t word) {
    int32_t sc = 0, best = 0, rd = 0;
    for (int32_t i = 0; i < word; ++i) {
        sc += M[Q[q + i] * kMatrixStride + S[s + i]];
        if (sc > best) {
            best = sc;
            rd = i + 1;
        }
    }
    const int32_t qr = q + rd, sr = s + rd;

    // left: positions sr-1-k, qr-1-k for k = 0..n
    const int32_t n = min(sr - 1, qr - 1);
    int32_t score = 0, mx = 0, left_len = 0;
    for (int32_t k = 0; k <= n; ++k) {
        score += M[Q[qr - 1 - k] * kMatrixStride + S[sr - 1 - k]];
        if (score > mx) {
            mx = score;
            left_len = k + 1;
        }
        if (mx - score >= xdrop) break;
    }
    TwoHit r;
    r.right_extend = left_len >= sr - s_left;
    r.s_last = sr;
    int32_t right_len = 0, total = mx;
    if (r.right_extend) {
        const int32_t n2 = min(slen - sr, qlen - qr);
        int32_t rscore = mx, rmax = mx, best_k = -1, k = 0;
        for (; k < n2; ++k) {
            rscore += M[Q[qr + k] * kMatrixStride + S[sr + k]];
            if (rscore > rmax) {
                rmax = rscore;
                best_k = k;
            }
            if (rscore <= 0 || rmax - rscore >= xdrop) break;
        }
        right_len = best_k + 1;
        r.s_last = sr + k;
        total = rmax;
    }
    r.hsp_q = qr - left_len;
    r.hsp_s = sr - left_len;
    r.len = left_len + right_len;
    r.score = total;
    return r;
}

template <int WORD, bool TABLE_IN_LDS>
__global__ __launch_bounds__(1024) void k_scan(ScanParams p) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid >> 6;
    const uint32_t nwaves = blockDim.x >> 6;

    // ---- stage the query side into LDS (once per workgroup) ----
    const uint32_t* table = p.table;
    if (TABLE_IN_LDS) {
        uint32_t* t = reinterpret_cast<uint32_t*>(lds + p.lds_table);
        for (uint32_t i = tid; i < p.table_words; i += blockDim.x) t[i] = p.table[i];
        table = t;
    }
    int8_t* M = reinterpret_cast<int8_t*>(lds + p.lds_matrix);
    for (uint32_t i = tid; i < 28 * kMatrixStride; i += blockDim.x) {
        const uint32_t r = i / kMatrixStride, c = i % kMatrixStride;
        M[i] = c < 28 ? p.matrix[r * 28 + c] : (int8_t)-128;
    }
    uint8_t* Q = lds + p.lds_query;
    for (int32_t i = (int32_t)tid; i < p.qlen; i += (int32_t)blockDim.x) Q[i] = p.query[i];

    const uint32_t dlen = (uint32_t)p.diag_mask + 1;
    uint32_t* diag = reinterpret_cast<uint32_t*>(lds + p.lds_wave + wave * p.wave_stride);
    uint32_t* claim = diag + dlen;
    uint32_t* ring = claim + dlen;
    const uint32_t qcap = p.qcap, qmask = qcap - 1;
    for (uint32_t i = lane; i < dlen; i += 64) {
        diag[i] = 0;
        claim[i] = 0xFFFFFFFFu;
    }
    __syncthreads();

    const uint2* groups = reinterpret_cast<const uint2*>(table);
    const uint32_t* starts = table + 2 * p.ngroups;
    const uint16_t* entries = reinterpret_cast<const uint16_t*>(starts + p.npresent + 1);
    const int32_t mask = p.diag_mask, window = p.window, xdrop = p.xdrop, cutoff = p.cutoff;

    const uint32_t wave_global = blockIdx.x * nwaves + wave;
    const uint32_t seg_base = wave_global * p.seg_cap;
    uint32_t seg_fill = 0;          // wave-uniform
    int32_t doff = window;          // NCBI diagonal offset, wave-uniform

    for (;;) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(p.work_counter, p.grab);
        item = uniform(__shfl(item, 0, 64));
        if (item >= p.norder) break;
        const uint32_t item_end = min(item + p.grab, p.norder);
        for (; item < item_end; ++item) {
            const uint32_t sid = p.order[item];
            const uint8_t* __restrict__ S = p.residues + p.soff[sid];
            const int32_t slen = (int32_t)p.slen[sid];
            const int32_t last = slen - WORD;  // last word start
            uint32_t hits = 0, extended = 0, saved = 0;
            const uint32_t rec_start = seg_fill;
            uint32_t head = 0, count = 0;  // ring state, wave-uniform
            int32_t qbase = 0;             // subject offset of ring entries' base

            // Applies the oldest n (<= 64) ring entries to the diagonal table.
            auto drain_batch = [&](uint32_t n) {
                bool active = lane < n;
                uint32_t e = 0;
                if (active) e = ring[(head + lane) & qmask];
                const int32_t q = (int32_t)(e >> 16);
                const int32_t s = qbase + (int32_t)(e & 0xFFFFu);
                const uint32_t d = (uint32_t)((q - s) & mask);
                while (__ballot(active)) {
                    if (active) atomicMin(&claim[d], lane);
                    wave_sync();
                    const bool win = active && claim[d] == lane;
                    wave_sync();
                    bool rec = false, ext = false;
                    TwoHit x;
                    if (win) {
                        claim[d] = 0xFFFFFFFFu;
                        const uint32_t st = diag[d];
                        const int32_t lastv = (int32_t)(st & ~kFlagBit);
                        uint32_t ns = st;
                        if (st & kFlagBit) {
                            if (s + doff >= lastv) ns = (uint32_t)(s + doff);
                        } else {
                            const int32_t lh = lastv - doff;
                            const int32_t diff = s - lh;
                            if (diff >= window) {
                                ns = (uint32_t)(s + doff);
                            } else if (diff >= WORD) {
                                ext = true;
                                x = extend_two_hit(M, Q, p.qlen, S, slen, lh + WORD, s, q, xdrop, WORD);
                                rec = x.score >= cutoff;
                                ns = x.right_extend ? ((uint32_t)(x.s_last - (WORD - 1) + doff) | kFlagBit)
                                                    : (uint32_t)(s + doff);
                            }
                        }
                        diag[d] = ns;
                        active = false;
                    }
                    extended += (uint32_t)__popcll(__ballot(ext));
                    const unsigned long long rb = __ballot(rec);
                    if (rb) {
                        const uint32_t slot = seg_fill + rank_below(rb);
                        if (rec) {
                            if (slot < p.seg_cap) {
                                bpx_init_hsp h;
                                h.hsp_q = x.hsp_q;
                                h.hsp_s = x.hsp_s;
                                h.query_offset = q;
                                h.subject_offset = s;
                                h.hsp_len = x.len;
                                h.score = x.score;
                                p.seg[seg_base + slot] = h;
                            } else {
                                atomicOr(p.overflow, 1u);
                            }
                        }
                        seg_fill += (uint32_t)__popcll(rb);
                        saved += (uint32_t)__popcll(rb);
                    }
                    wave_sync();
                }
                head = (head + n) & qmask;
                count -= n;
            };
            auto drain_all = [&]() {
                while (count > 0) drain_batch(min(count, 64u));
            };

            for (int32_t pos = 0; pos <= last;) {
                if (count == 0) {
                    qbase = pos;
                } else if (pos + 64 - qbase >= 65536) {
                    drain_all();
                    qbase = pos;
                }
                const int32_t s = pos + (int32_t)lane;
                uint32_t cnt = 0, start = 0;
                if (s <= last) {
                    uint32_t w = S[s];
#pragma unroll
                    for (int i = 1; i < WORD; ++i) w = (w << 5) | S[s + i];
                    const uint2 g = groups[w >> 5];
                    const uint32_t bit = 1u << (w & 31u);
                    if (g.x & bit) {
                        const uint32_t r = g.y + (uint32_t)__popc(g.x & (bit - 1u));
                        start = starts[r];
                        cnt = starts[r + 1] - start;
                    }
                }
                const uint32_t incl = wave_inclusive_sum(cnt, lane);
                const bool fits = incl <= qcap - count;
                const unsigned long long fb = __ballot(fits);
                const uint32_t nfit = (uint32_t)__popcll(fb);
                if (nfit == 0) {
                    if (count == 0) {  // a cell larger than the ring: host sizing error
                        if (lane == 0) atomicOr(p.overflow, 2u);
                        break;
                    }
                    drain_all();
                    continue;
                }
                if (fits && cnt) {
                    const uint32_t at = head + count + (incl - cnt);
                    const uint32_t sv = (uint32_t)(s - qbase);
                    for (uint32_t j = 0; j < cnt; ++j)
                        ring[(at + j) & qmask] = ((uint32_t)entries[start + j] << 16) | sv;
                }
                const uint32_t added = uniform(__shfl(incl, (int)nfit - 1, 64));
                count += added;
                hits += added;
                pos += (int32_t)nfit;
                wave_sync();
                while (count >= 64) drain_batch(64);
            }
            drain_all();

            if (lane == 0) {
                const uint32_t slot = sid - p.first;
                p.stats[3 * (size_t)slot + 0] = (int32_t)hits;
                p.stats[3 * (size_t)slot + 1] = (int32_t)saved;
                p.stats[3 * (size_t)slot + 2] = (int32_t)extended;
                p.rec_loc[slot] = seg_base + rec_start;
            }
            // Blast_ExtendWordExit: move to a fresh diagonal range, clearing
            // the table before the 31-bit field can run out.
            if (doff >= 0x7FFFFFFF / 4) {
                for (uint32_t i = lane; i < dlen; i += 64) diag[i] = 0;
                doff = window;
                wave_sync();
            } else {
                doff += slen + window;
            }
        }
    }
}

// ---- compaction -------------------------------------------------------------

constexpr uint32_t kCompactThreads = 256;
constexpr uint32_t kPer = kTile / kCompactThreads;  // 4 subjects per thread

__device__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* sh, uint32_t* total) {
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v, lane);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
        if (w < wave) base += sh[w];
        all += sh[w];
    }
    __syncthreads();
    *total = all;
    return base + incl - v;
}

// Per tile: sum of saved counts, plus the query-level ungapped statistics
// Blast_UngappedStatsUpdate accumulates (blast_engine.c:553-556).
__global__ __launch_bounds__(kCompactThreads) void k_tile_sums(CompactParams p) {
    __shared__ uint32_t sh[kCompactThreads / 64];
    __shared__ unsigned long long acc[5];
    if (threadIdx.x < 5) acc[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t sum = 0;
    unsigned long long h = 0, nh = 0, ex = 0, sv = 0, np = 0;
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = base + k;
        if (i < p.count) {
            const int32_t* st = p.stats + 3 * (size_t)i;
            sum += (uint32_t)st[1];
            if (st[0] > 0) {  // Blast_UngappedStatsUpdate ignores subjects without hits
                h += (uint32_t)st[0];
                nh += 1;
                ex += (uint32_t)st[2];
                sv += (uint32_t)st[1];
                np += st[1] > 0;
            }
        }
    }
    uint32_t tot;
    block_exclusive_sum(sum, sh, &tot);
    if (threadIdx.x == 0) p.tile_sum[blockIdx.x] = tot;
    atomicAdd(&acc[0], h);
    atomicAdd(&acc[1], nh);
    atomicAdd(&acc[2], ex);
    atomicAdd(&acc[3], sv);
    atomicAdd(&acc[4], np);
    __syncthreads();
    if (threadIdx.x < 5 && acc[threadIdx.x]) atomicAdd(&p.ustats[threadIdx.x], acc[threadIdx.x]);
}

// Exclusive scan of the tile sums in one workgroup (tile offsets in place).
__global__ __launch_bounds__(1024) void k_tile_scan(CompactParams p, uint32_t ntiles) {
    __shared__ uint32_t sh[16];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < ntiles; b += blockDim.x) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < ntiles ? p.tile_sum[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_sum(v, sh, &tot);
        if (i < ntiles) p.tile_sum[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        p.rec_off[p.count] = carry;
        p.ustats[5] = carry;
    }
}

// Per subject: CSR offset, then copy its records from the wave segment in
// (subject_offset, query_offset) order -- the sequential emission order.
__global__ __launch_bounds__(kCompactThreads) void k_gather(CompactParams p) {
    __shared__ uint32_t sh[kCompactThreads / 64];
    if (*p.overflow) return;  // segments incomplete; the host rescans
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t c[kPer], sum = 0;
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = base + k;
        c[k] = i < p.count ? (uint32_t)p.stats[3 * (size_t)i + 1] : 0u;
        sum += c[k];
    }
    uint32_t tot;
    uint32_t off = p.tile_sum[blockIdx.x] + block_exclusive_sum(sum, sh, &tot);
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = base + k;
        if (i >= p.count) break;
        p.rec_off[i] = off;
        const bpx_init_hsp* src = p.seg + p.rec_loc[i];
        bpx_init_hsp* dst = p.records + off;
        for (uint32_t j = 0; j < c[k]; ++j) {
            const bpx_init_hsp h = src[j];
            uint32_t m = j;  // insertion sort on (subject_offset, query_offset)
            while (m > 0) {
                const bpx_init_hsp g = dst[m - 1];
                if (g.subject_offset < h.subject_offset ||
                    (g.subject_offset == h.subject_offset && g.query_offset < h.query_offset))
                    break;
                dst[m] = g;
                --m;
            }
            dst[m] = h;
            p.subjects[off + j] = p.first + i;
        }
        off += c[k];
    }
}

__global__ void k_fill_u32(uint32_t* dst, uint32_t value, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

}  // namespace

LdsPlan plan_lds(uint32_t table_bytes, int32_t qlen, int32_t diag_len, uint32_t qcap, int waves,
                 uint32_t lds_limit, bool allow_table_in_lds) {
    auto up16 = [](uint32_t v) { return (v + 15u) & ~15u; };
    LdsPlan pl{};
    pl.wave_stride = up16((uint32_t)diag_len * 8u + qcap * 4u);
    const uint32_t fixed = up16(28 * kMatrixStride) + up16((uint32_t)qlen) + pl.wave_stride * (uint32_t)waves;
    pl.table_in_lds = allow_table_in_lds && up16(table_bytes) + fixed <= lds_limit;
    uint32_t at = 0;
    pl.table = at;
    if (pl.table_in_lds) at += up16(table_bytes);
    pl.matrix = at;
    at += up16(28 * kMatrixStride);
    pl.query = at;
    at += up16((uint32_t)qlen);
    pl.wave = at;
    at += pl.wave_stride * (uint32_t)waves;
    pl.total = at;
    return pl;
}

hipError_t launch_scan(const ScanParams& p, bool table_in_lds, int waves, int groups,
                       hipStream_t stream) {
    const dim3 grid((unsigned)groups), block((unsigned)(64 * waves));
    void (*fn)(ScanParams) = nullptr;
    if (p.word == 3)
        fn = table_in_lds ? k_scan<3, true> : k_scan<3, false>;
    else if (p.word == 2)
        fn = table_in_lds ? k_scan<2, true> : k_scan<2, false>;
    else
        return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds_total);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, grid, block, p.lds_total, stream, p);
    return hipGetLastError();
}

hipError_t launch_compact(const CompactParams& p, hipStream_t stream) {
    const uint32_t ntiles = (p.count + kTile - 1) / kTile;
    if (ntiles == 0) {
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, stream, p, 0u);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_tile_sums, dim3(ntiles), dim3(kCompactThreads), 0, stream, p);
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, stream, p, ntiles);
    hipLaunchKernelGGL(k_gather, dim3(ntiles), dim3(kCompactThreads), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* dst, uint32_t value, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0,
                       stream, dst, value, n);
    return hipGetLastError();
}

}  // namespace bpx
