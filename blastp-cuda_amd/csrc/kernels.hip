// kernels.hip -- gfx950 kernels of the blastp word finder.
//
// k_scan: two-hit word finding + ungapped X-drop extension of a group of
// queries against a list of subjects.  The reference (gpu_blastp_kernel.cu:
// 42-320) gives each CUDA thread a whole subject with a private diagonal
// array in global memory, so every word hit is a random global read-modify-
// write and every extension step a dependent global load.  Here a 64-lane
// wavefront owns a pack of subjects at a time -- one long subject, or up to
// four consecutive ones that fit its subject buffer together, scanned as one
// stream (each at NCBI's running diagonal offset from the previous one, so
// the two-hit state is exactly the per-subject one):
//
//   * the pack is staged into a per-wave LDS buffer with 16-byte loads
//     (long subjects through a sliding window); words and extensions read it
//     from LDS;
//   * the workgroup stages the group's packed lookup table, its query
//     residues and a 28x32 score table into LDS once per group;
//   * the lanes take 64 subject positions at a time: each forms its word,
//     tests the presence bits and finds its cell; a DPP prefix sum over the
//     cell sizes compacts the hits, in (subject offset, query offset) order,
//     into a per-wave LDS ring.  In steady state the lookups of the next 64
//     positions run interleaved with the claim rounds of the current batch;
//   * the ring is drained 64 hits at a time.  Hits that share a diagonal
//     slot are applied in scan order by claiming a hashed slot with an LDS
//     atomic-min on the lane number (gpu_blastp_kernel.cu:194-301 applied
//     exactly, in parallel across slots); hits left after two claim rounds
//     return to the front of the ring for the next batch;
//   * a hit that completes a pair does not extend inline: its slot is marked
//     pending and the extension is queued.  A later hit on a pending slot is
//     deferred, in order, and applied again after the queue is flushed, so
//     every slot still sees its hits strictly in scan order.  Queued
//     extensions (s_BlastAaExtendTwoHit/Left/Right semantics,
//     gpu_blastp_kernel.cu:341-549) run when a queue fills or the pack
//     ends: one per lane, or a few on 8-lane segments;
//   * records go to the wave's private output segment, tagged with their
//     subject; per-(query, subject) counts, optional counters and per-query
//     ungapped statistics are written when the pack is done.
//
// Records are put in (query, subject, subject offset, query offset) order by
// the count / scan / gather kernels below (no atomics on the output).
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace bpx {

namespace {

// Diagonal slot state {value, reset flag, pending}.  Wide slots (u32) hold
// last_hit + NCBI's running diagonal offset; narrow slots (u16) hold
// last_hit + window for the current subject only (the table is cleared per
// subject), so a zero slot reads as "no hit yet" in both.
template <bool NARROW>
struct Slot {
    using T = uint32_t;
    static constexpr uint32_t kFlag = 0x80000000u;  // NCBI reset flag: an extension reached past this hit
    static constexpr uint32_t kPend = 0x40000000u;  // extension queued, outcome not yet applied
    static constexpr uint32_t kVal = 0x3FFFFFFFu;
};
template <>
struct Slot<true> {
    using T = uint16_t;
    static constexpr uint32_t kFlag = 0x8000u;
    static constexpr uint32_t kPend = 0x4000u;
    static constexpr uint32_t kVal = 0x3FFFu;
};
constexpr int32_t kOffsetLimit = 1 << 28;
constexpr uint32_t kSeg = 8;       // lanes per cooperative extension task
constexpr uint32_t kCoopMax = 16;  // flushes of at most this many tasks (of a resident subject) run cooperatively
constexpr int kMatrixStride = 32;        // LDS score rows padded 28 -> 32
constexpr int kMatrixRows = 32;          // ... and 28 -> 32 rows: any 10-bit index is inside
constexpr uint32_t kCtlBytes = 16;       // workgroup control words at LDS offset 0

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t rank_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Inclusive prefix sum over the 64 lanes with DPP: Hillis-Steele inside each
// 16-lane row (row_shr 1, 2, 4, 8), then row 0's total into row 1 and row 2's
// into row 3 (row_bcast:15), then lane 31's total into rows 2-3
// (row_bcast:31).
// bound_ctrl reads 0 for a lane whose source is outside the row, so each
// step folds into one v_add_u32 with a DPP source operand.  The empty asm
// hides the partial sums from the caller's arithmetic: otherwise the
// compiler rebuilds incl - v from them and keeps the separate moves.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, true);
    asm volatile("" : "+v"(v));
    return v;
}

struct TwoHit {
    int32_t hsp_q, hsp_s, len, score, s_last;
    bool right_extend;
};

// Subject residue p of the wave's staged subject.  sbuf holds positions
// [wlo, whi) (all of a subject that fits); any other position -- behind the
// window, or ahead of it when a right extension runs past the loaded part --
// comes from HBM.
__device__ __forceinline__ uint32_t sres(const uint8_t* sbuf, uint32_t smask, const uint8_t* S, int32_t p,
                                         int32_t wlo, int32_t whi, int32_t slen) {
    uint32_t v = sbuf[p & smask];
    if (p < wlo || p >= whi) v = (p >= 0 && p < slen) ? S[p] : 0u;
    return v;
}

// s_BlastAaExtendTwoHit semantics (gpu_blastp_kernel.cu:341-415 with the
// loops of :435-481 and :503-549): best prefix of the second word, X-drop to
// the left from one past it, to the right only if the left reached the
// first hit.  c is the second hit's position in the group's concatenated
// residues Q (qcap bytes); the query occupies Q[qstart, qstart + qlen).
//
// Both X-drop loops run in blocks of kBlock steps: the block's residues and
// scores are loaded together, then the steps are applied in order in
// registers under a `done` predicate, so each step sees exactly the
// sequential running score, maximum and stop test.  Two steps per block
// measured fastest: the lanes of a flush end their walks at different
// steps, and every step a block runs past a lane's end is issued for the
// whole wave.  RESIDENT (the whole
// subject is in sbuf): every load is an unconditional LDS read at a clamped
// index, masked afterwards.
constexpr int kBlock = 2;

template <bool RESIDENT>
__device__ __forceinline__ uint32_t sread(const uint8_t* sbuf, uint32_t smask, const uint8_t* S, int32_t p,
                                          int32_t wlo, int32_t whi, int32_t slen) {
    if (RESIDENT) return sbuf[p & smask];
    return sres(sbuf, smask, S, p, wlo, whi, slen);
}

template <int WORD, bool RESIDENT>
__device__ __forceinline__ TwoHit extend_two_hit(const int8_t* M, const uint8_t* Q, uint32_t qcap, int32_t qstart,
                                                 int32_t qlen, const uint8_t* sbuf, uint32_t smask,
                                                 const uint8_t* S, int32_t wlo, int32_t whi, int32_t slen,
                                                 int32_t s_left, int32_t s, int32_t c, int32_t xdrop) {
    int32_t sc = 0, best = 0, rd = 0;
#pragma unroll
    for (int32_t i = 0; i < WORD; ++i) {
        sc += M[Q[c + i] * kMatrixStride + sread<RESIDENT>(sbuf, smask, S, s + i, wlo, whi, slen)];
        if (sc > best) {
            best = sc;
            rd = i + 1;
        }
    }
    const int32_t cr = c + rd, sr = s + rd, qr = cr - qstart;

    // left: positions (cr-1-k, sr-1-k) for k = 0..n
    const int32_t n = min(sr - 1, qr - 1);
    int32_t score = 0, mx = 0, left_len = 0;
    bool done = n < 0;
    for (int32_t k0 = 0; !done; k0 += kBlock) {
        int32_t m[kBlock];
        uint32_t qb[kBlock], sb[kBlock];
#pragma unroll
        for (int i = 0; i < kBlock; ++i) {
            const int32_t k = k0 + i;
            qb[i] = Q[max(cr - 1 - k, 0)];
            sb[i] = RESIDENT ? (uint32_t)sbuf[max(sr - 1 - k, 0) & smask]
                             : (k <= n ? sres(sbuf, smask, S, sr - 1 - k, wlo, whi, slen) : 0u);
        }
#pragma unroll
        for (int i = 0; i < kBlock; ++i) m[i] = M[qb[i] * kMatrixStride + sb[i]];
#pragma unroll
        for (int i = 0; i < kBlock; ++i) {
            const int32_t k = k0 + i;
            const bool ok = !done && k <= n;
            score = ok ? score + m[i] : score;
            const bool up = ok && score > mx;
            mx = up ? score : mx;
            left_len = up ? k + 1 : left_len;
            done = done || k >= n || (ok && mx - score >= xdrop);
        }
    }
    TwoHit r;
    r.right_extend = left_len >= sr - s_left;
    r.s_last = sr;
    int32_t right_len = 0, total = mx;
    if (r.right_extend) {
        // right: positions (cr+k, sr+k) for k = 0..n2-1, from the left score
        const int32_t n2 = min(slen - sr, qlen - qr);
        int32_t rscore = mx, rmax = mx, best_k = -1, kstop = n2;
        bool rdone = n2 <= 0;
        for (int32_t k0 = 0; !rdone; k0 += kBlock) {
            int32_t m[kBlock];
            uint32_t qb[kBlock], sb[kBlock];
#pragma unroll
            for (int i = 0; i < kBlock; ++i) {
                const int32_t k = k0 + i;
                qb[i] = Q[min((uint32_t)(cr + k), qcap - 1)];
                sb[i] = RESIDENT ? (uint32_t)sbuf[(sr + k) & smask]
                                 : (k < n2 ? sres(sbuf, smask, S, sr + k, wlo, whi, slen) : 0u);
            }
#pragma unroll
            for (int i = 0; i < kBlock; ++i) m[i] = M[qb[i] * kMatrixStride + sb[i]];
#pragma unroll
            for (int i = 0; i < kBlock; ++i) {
                const int32_t k = k0 + i;
                const bool ok = !rdone && k < n2;
                rscore = ok ? rscore + m[i] : rscore;
                const bool up = ok && rscore > rmax;
                rmax = up ? rscore : rmax;
                best_k = up ? k : best_k;
                const bool stop = ok && (rscore <= 0 || rmax - rscore >= xdrop);
                kstop = stop ? k : kstop;
                rdone = rdone || stop || k + 1 >= n2;
            }
        }
        right_len = best_k + 1;
        r.s_last = sr + kstop;
        total = rmax;
    }
    r.hsp_q = qr - left_len;
    r.hsp_s = sr - left_len;
    r.len = left_len + right_len;
    r.score = total;
    return r;
}

// extend_two_hit for a subject resident in sbuf, software-pipelined: each
// iteration applies two steps whose scores were loaded one iteration
// earlier, while the next two steps' scores and the residues of the two
// after them are in flight, so a lane waits on no load of its own
// iteration.  Loads run up to five steps past a walk's end: subject reads
// wrap inside sbuf, query reads stay within six bytes of the group's
// residues (inside the LDS allocation); their values are never used.  (A
// general form with two, three or four steps per iteration measured slower.)
template <int WORD>
__device__ __forceinline__ TwoHit extend_resident(const int8_t* M, const uint8_t* Q, int32_t qstart, int32_t qlen,
                                                  const uint8_t* sbuf, uint32_t smask, uint32_t sbase, int32_t slen,
                                                  int32_t s_left, int32_t s, int32_t c, int32_t xdrop) {
    int32_t sc = 0, best = 0, rd = 0;
#pragma unroll
    for (int32_t i = 0; i < WORD; ++i) {
        sc += M[Q[c + i] * kMatrixStride + sbuf[(sbase + (uint32_t)(s + i)) & smask]];
        if (sc > best) {
            best = sc;
            rd = i + 1;
        }
    }
    const int32_t cr = c + rd, sr = s + rd, qr = cr - qstart;
    // residues past a walk's end may hold any byte: the index stays inside
    // the 32 x 32 LDS score table (a residue code is below 28)
    auto score_at = [&](uint32_t qb, uint32_t sb) { return (int32_t)M[((qb << 5) | sb) & 1023u]; };

    // left: step k at (cr-1-k, sr-1-k), k = 0..n
    const int32_t n = min(sr - 1, qr - 1);
    int32_t mx = 0, left_len = 0;
    if (n >= 0) {
        const uint8_t* qp = Q + (cr - 1);
        const uint32_t sp = sbase + (uint32_t)(sr - 1);
        int32_t m0 = score_at(qp[0], sbuf[sp & smask]);
        int32_t m1 = score_at(qp[-1], sbuf[(sp - 1u) & smask]);
        uint32_t q2 = qp[-2], q3 = qp[-3], s2 = sbuf[(sp - 2u) & smask], s3 = sbuf[(sp - 3u) & smask];
        int32_t score = 0;
        bool done = false;
        for (int32_t k0 = 0; !done; k0 += 2) {
            const int32_t n0 = score_at(q2, s2), n1 = score_at(q3, s3);
            q2 = qp[-(k0 + 4)];
            q3 = qp[-(k0 + 5)];
            s2 = sbuf[(sp - (uint32_t)(k0 + 4)) & smask];
            s3 = sbuf[(sp - (uint32_t)(k0 + 5)) & smask];
            // step k0 (always inside the walk), then k0 + 1 unless it ended
            const int32_t a = score + m0;
            const bool up0 = a > mx;
            mx = up0 ? a : mx;
            left_len = up0 ? k0 + 1 : left_len;
            const bool end0 = k0 >= n || mx - a >= xdrop;
            const int32_t b = a + m1;
            const bool up1 = !end0 && b > mx;
            mx = up1 ? b : mx;
            left_len = up1 ? k0 + 2 : left_len;
            done = end0 || k0 + 1 >= n || mx - b >= xdrop;
            score = b;
            m0 = n0;
            m1 = n1;
        }
    }
    TwoHit r;
    r.right_extend = left_len >= sr - s_left;
    r.s_last = sr;
    int32_t right_len = 0, total = mx;
    if (r.right_extend) {
        // right: step k at (cr+k, sr+k), k = 0..n2-1, from the left score
        const int32_t n2 = min(slen - sr, qlen - qr);
        int32_t rmax = mx, best_k = -1, kstop = n2;
        if (n2 > 0) {
            const uint8_t* qp = Q + cr;
            const uint32_t sp = sbase + (uint32_t)sr;
            int32_t m0 = score_at(qp[0], sbuf[sp & smask]);
            int32_t m1 = score_at(qp[1], sbuf[(sp + 1u) & smask]);
            uint32_t q2 = qp[2], q3 = qp[3], s2 = sbuf[(sp + 2u) & smask], s3 = sbuf[(sp + 3u) & smask];
            int32_t rscore = mx;
            bool done = false;
            for (int32_t k0 = 0; !done; k0 += 2) {
                const int32_t n0 = score_at(q2, s2), n1 = score_at(q3, s3);
                q2 = qp[k0 + 4];
                q3 = qp[k0 + 5];
                s2 = sbuf[(sp + (uint32_t)(k0 + 4)) & smask];
                s3 = sbuf[(sp + (uint32_t)(k0 + 5)) & smask];
                const int32_t a = rscore + m0;
                const bool up0 = a > rmax;
                rmax = up0 ? a : rmax;
                best_k = up0 ? k0 : best_k;
                const bool stop0 = a <= 0 || rmax - a >= xdrop;
                const bool end0 = stop0 || k0 + 1 >= n2;
                const int32_t b = a + m1;
                const bool up1 = !end0 && b > rmax;
                rmax = up1 ? b : rmax;
                best_k = up1 ? k0 + 1 : best_k;
                const bool stop1 = !end0 && (b <= 0 || rmax - b >= xdrop);
                kstop = stop0 ? k0 : (stop1 ? k0 + 1 : kstop);
                done = end0 || stop1 || k0 + 2 >= n2;
                rscore = b;
                m0 = n0;
                m1 = n1;
            }
        }
        right_len = best_k + 1;
        r.s_last = sr + kstop;
        total = rmax;
    }
    r.hsp_q = qr - left_len;
    r.hsp_s = sr - left_len;
    r.len = left_len + right_len;
    r.score = total;
    return r;
}

// Inclusive prefix sum / maximum inside each W-lane segment (W = 8 or 16) of
// the wave: DPP row_shr 1, 2, 4 (and 8 for W = 16); a lane whose source lies
// in the previous segment (or outside the 16-lane row) takes the identity.
template <int W>
__device__ __forceinline__ int32_t seg_incl_sum(int32_t v, uint32_t rl) {
    int32_t t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
    v += rl >= 1 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
    v += rl >= 2 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
    v += rl >= 4 ? t : 0;
    if (W == 16) {
        t = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
        v += rl >= 8 ? t : 0;
    }
    return v;
}
template <int W>
__device__ __forceinline__ int32_t seg_incl_max(int32_t v, uint32_t rl) {
    constexpr int32_t lo = (int32_t)0x80000000;
    int32_t t;
    t = __builtin_amdgcn_update_dpp(lo, v, 0x111, 0xf, 0xf, false);
    v = rl >= 1 ? max(v, t) : v;
    t = __builtin_amdgcn_update_dpp(lo, v, 0x112, 0xf, 0xf, false);
    v = rl >= 2 ? max(v, t) : v;
    t = __builtin_amdgcn_update_dpp(lo, v, 0x114, 0xf, 0xf, false);
    v = rl >= 4 ? max(v, t) : v;
    if (W == 16) {
        t = __builtin_amdgcn_update_dpp(lo, v, 0x118, 0xf, 0xf, false);
        v = rl >= 8 ? max(v, t) : v;
    }
    return v;
}

// One X-drop pass of a W-lane segment: W consecutive steps per iteration.
// Step k of the segment has score m_k; the running score is carry + prefix
// sum, the running maximum the prefix maximum, and the first lane whose stop
// test holds (or the first step past the end) ends the pass -- the
// sequential loop of gpu_blastp_kernel.cu:451-474 (left) / :523-541 (right),
// evaluated W steps at a time.  RIGHT adds the stop on a non-positive running
// score.  Returns (via refs) the maximum, the step of its first occurrence
// (-1 if it never rose above the start value) and the stop step.
template <bool RIGHT, int W>
__device__ __forceinline__ void seg_xdrop(const int8_t* M, const uint8_t* Q, uint32_t qcap, const uint8_t* sbuf,
                                          uint32_t smask, uint32_t sbase, int32_t cq, int32_t cs, int32_t nsteps,
                                          int32_t start,
                                          int32_t xdrop, uint32_t lane, int32_t& out_max, int32_t& out_best,
                                          int32_t& out_stop) {
    constexpr int32_t kNeg = -(1 << 22);
    constexpr uint32_t kSegMask = (1u << W) - 1u;
    const uint32_t rl = lane & (W - 1u), rbase = lane & ~(W - 1u);
    int32_t carry = start, mx = start, best = -1, stop_at = nsteps;
    bool done = nsteps <= 0;
    // left steps walk down from (cq, cs), right steps walk up from it.  The
    // loads are pipelined: an iteration uses the score loaded one iteration
    // earlier and the residues loaded two earlier (query reads clamped to the
    // group's residues, subject reads wrapped in sbuf, score index kept in
    // the 32 x 32 table: values past the walk are masked, never used)
    auto q_at = [&](int32_t k) { return (uint32_t)Q[min((uint32_t)max(RIGHT ? cq + k : cq - k, 0), qcap - 1)]; };
    auto s_at = [&](int32_t k) { return (uint32_t)sbuf[(sbase + (uint32_t)(RIGHT ? cs + k : cs - k)) & smask]; };
    auto m_at = [&](uint32_t qb, uint32_t sb) { return (int32_t)M[((qb << 5) | sb) & 1023u]; };
    int32_t m_next = m_at(q_at((int32_t)rl), s_at((int32_t)rl));
    uint32_t q2 = q_at(W + (int32_t)rl), s2 = s_at(W + (int32_t)rl);
    for (int32_t k0 = 0; __ballot(!done) != 0; k0 += W) {
        const int32_t k = k0 + (int32_t)rl;
        const bool ok = !done && k < nsteps;
        const int32_t m_raw = m_next;
        m_next = m_at(q2, s2);
        q2 = q_at(k + 2 * W);
        s2 = s_at(k + 2 * W);
        const int32_t m = ok ? m_raw : kNeg;
        const int32_t run = carry + seg_incl_sum<W>(m, rl);
        const int32_t pm = max(mx, seg_incl_max<W>(run, rl));
        const bool stop = ok ? ((RIGHT && run <= 0) || pm - run >= xdrop) : true;
        const uint32_t smaskr = (uint32_t)(__ballot(stop) >> rbase) & kSegMask;
        const uint32_t kb = smaskr ? (uint32_t)__ffs(smaskr) - 1u : W - 1u;
        const int32_t pm_kb = __shfl(pm, (int)(rbase + kb), 64);
        const int32_t run_last = __shfl(run, (int)(rbase + W - 1u), 64);
        const uint32_t eq = (uint32_t)(__ballot(run == pm_kb && rl <= kb) >> rbase) & kSegMask;
        if (!done) {
            if (pm_kb > mx) {
                mx = pm_kb;
                best = k0 + (int32_t)__ffs(eq) - 1;
            }
            carry = run_last;
            if (smaskr) {
                done = true;
                stop_at = k0 + (int32_t)kb;
            }
        }
    }
    out_max = mx;
    out_best = best;
    out_stop = stop_at;
}

// s_BlastAaExtendTwoHit for one task per W-lane segment (all 64 lanes call
// it; a segment with valid == false does no work).  Same result as
// extend_two_hit for a subject resident in sbuf.
template <int WORD, int W>
__device__ __forceinline__ TwoHit coop_extend_two_hit(const int8_t* M, const uint8_t* Q, uint32_t qcap,
                                                      const uint8_t* sbuf, uint32_t smask, uint32_t sbase,
                                                      int32_t slen, bool valid,
                                                      int32_t qstart, int32_t qlen, int32_t s_left, int32_t s,
                                                      int32_t c, int32_t xdrop, uint32_t lane) {
    int32_t sc = 0, best = 0, rd = 0;
#pragma unroll
    for (int32_t i = 0; i < WORD; ++i) {
        sc += M[Q[min((uint32_t)(c + i), qcap - 1)] * kMatrixStride + sbuf[(sbase + (uint32_t)(s + i)) & smask]];
        if (sc > best) {
            best = sc;
            rd = i + 1;
        }
    }
    const int32_t cr = c + rd, sr = s + rd, qr = cr - qstart;
    // left: steps k = 0..n at (cr-1-k, sr-1-k)
    const int32_t nl = valid ? min(sr - 1, qr - 1) + 1 : 0;
    int32_t lmax, lbest, lstop;
    seg_xdrop<false, W>(M, Q, qcap, sbuf, smask, sbase, cr - 1, sr - 1, nl, 0, xdrop, lane, lmax, lbest, lstop);
    const int32_t left_len = lbest + 1;
    TwoHit r;
    r.right_extend = valid && left_len >= sr - s_left;
    const int32_t nr = r.right_extend ? min(slen - sr, qlen - qr) : 0;
    int32_t rmax, rbest, rstop;
    seg_xdrop<true, W>(M, Q, qcap, sbuf, smask, sbase, cr, sr, nr, lmax, xdrop, lane, rmax, rbest, rstop);
    r.s_last = r.right_extend ? sr + rstop : sr;
    r.hsp_q = qr - left_len;
    r.hsp_s = sr - left_len;
    r.len = left_len + (r.right_extend ? rbest + 1 : 0);
    r.score = r.right_extend ? rmax : lmax;
    return r;
}

// Per-query counters of a wave (MULTI groups), 16 bytes each.
struct QCount {
    uint32_t hits, ext, saved, pad;
};
// A wavefront scans a pack of up to kPack consecutive subjects that fit its
// subject buffer together as one stream (see the pack loop in k_scan).
constexpr int kPack = 4;
// Per-subject counters of a pack (single-query groups).
struct PackSub {
    uint32_t hits, ext, saved, pad;
};
// Per-query ungapped statistics of a work item (MULTI groups) live in the
// registers of the lane that owns the query (lane j owns queries j, j + 64,
// j + 128, j + 192), not in LDS: a smaller per-wave LDS share leaves room
// for more waves.
constexpr int kQuerySlots = 4;  // queries per lane: groups hold at most 255

template <int WORD, bool TAB_LDS, bool MULTI, bool NARROW>
__global__ __launch_bounds__(64 * kMaxWaves) void k_scan(ScanParams p) {
    using SlotT = typename Slot<NARROW>::T;
    constexpr uint32_t kFlag = Slot<NARROW>::kFlag, kPend = Slot<NARROW>::kPend, kVal = Slot<NARROW>::kVal;
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid >> 6;
    const uint32_t nwaves = blockDim.x >> 6;
    uint32_t* ctl = reinterpret_cast<uint32_t*>(lds);

    int8_t* M = reinterpret_cast<int8_t*>(lds + p.lds_matrix);
    for (uint32_t i = tid; i < kMatrixRows * kMatrixStride; i += blockDim.x) {
        const uint32_t r = i / kMatrixStride, c = i % kMatrixStride;
        M[i] = (r < 28 && c < 28) ? p.matrix[r * 28 + c] : (int8_t)-128;
    }
    uint32_t* tab_lds = reinterpret_cast<uint32_t*>(lds + p.lds_table);
    uint8_t* Q = lds + p.lds_qres;
    uint8_t* qmap = lds + p.lds_qmap;
    QueryInfo* qinfo = reinterpret_cast<QueryInfo*>(lds + p.lds_qinfo);

    uint8_t* wbase = lds + p.lds_wave + wave * p.wave_stride;
    uint8_t* sbuf = wbase;
    const uint32_t diag_bytes = (p.diag_cap * (uint32_t)sizeof(SlotT) + 15u) & ~15u;
    SlotT* diag = reinterpret_cast<SlotT*>(wbase + p.sbuf);
    uint32_t* claim = reinterpret_cast<uint32_t*>(wbase + p.sbuf + diag_bytes);
    uint32_t* ring = claim + p.claim_slots;
    // queued extension: {slot | second hit's group position << 16, subject
    // offset - qbase (16 bits: qbase does not move while extensions are
    // queued)}; the first hit stays in the pending slot's value bits
    uint32_t* t_sc = ring + p.qcap;
    uint16_t* t_s = reinterpret_cast<uint16_t*>(t_sc + kTaskCap);  // subject offset - qbase
    uint32_t* dq = reinterpret_cast<uint32_t*>(t_s + kTaskCap);
    PackSub* pk = reinterpret_cast<PackSub*>(dq + kDefCap);  // this pack's subjects
    QCount* qc = reinterpret_cast<QCount*>(pk + kPack);        // MULTI: kPack x max_nq
    // zero-fills the wave's diagonal table, 16 bytes per lane and store
    auto clear_diag = [&]() {
        uint4* d4 = reinterpret_cast<uint4*>(diag);
        for (uint32_t i = lane; i < diag_bytes / 16u; i += 64) d4[i] = uint4{0u, 0u, 0u, 0u};
    };
    clear_diag();
    for (uint32_t i = lane; i < p.claim_slots; i += 64) claim[i] = 0xFFFFFFFFu;
    if (MULTI)
        for (uint32_t i = lane; i < p.max_nq * kPack; i += 64) qc[i] = QCount{};
    if (lane < kPack) pk[lane] = PackSub{};

    const uint32_t smask = p.sbuf - 1;
    const uint32_t qcap = p.qcap, qmask = qcap - 1;
    const int32_t window = p.window;
    const uint32_t seg_base = (p.seg_wave_base + blockIdx.x * nwaves + wave) * p.seg_cap;
    uint32_t seg_fill = 0;      // wave-uniform
    // optional per-wave phase times (p.phase), on the 100 MHz real-time clock:
    // lifetime, staging of subjects HBM -> LDS, group changes (barriers and
    // restaging), extension flushes (with the number of extensions run)
    const bool timed = p.phase != nullptr;
    const uint64_t t_begin = timed ? __builtin_amdgcn_s_memrealtime() : 0;
    uint64_t t_stage = 0, t_group = 0, n_staged = 0, b_staged = 0, t_flush = 0, n_flushed = 0;
    int32_t doff = window;      // running NCBI diagonal offset, wave-uniform
    uint32_t cur_group = 0xFFFFFFFFu;
    GroupInfo g{};

    // per-query ungapped statistics (Blast_UngappedStatsUpdate), summed over
    // the items this wave ran of one group -- a group's cost bins are
    // consecutive items -- and added to the launch's totals when the group
    // changes (flat) or the item ends (claimed items); lane j owns the
    // group's queries j, j + 64, ...
    uint32_t a_nseq = 0, a_ext = 0, a_saved = 0, a_passed = 0;  // single-query aggregates
    unsigned long long a_hits = 0;
    unsigned long long m_hits[kQuerySlots];                        // MULTI: per owned query
    uint32_t m_nseq[kQuerySlots], m_ext[kQuerySlots], m_saved[kQuerySlots], m_passed[kQuerySlots];
    auto zero_ustats = [&]() {
        a_nseq = a_ext = a_saved = a_passed = 0;
        a_hits = 0;
#pragma unroll
        for (int t = 0; t < kQuerySlots; ++t) {
            m_hits[t] = 0;
            m_nseq[t] = m_ext[t] = m_saved[t] = m_passed[t] = 0;
        }
    };
    // reads the current group's qinfo: called before the group changes
    auto flush_ustats = [&]() {
        if (!MULTI) {
            if (lane == 0 && a_nseq) {
                unsigned long long* u = p.ustats + 5 * (size_t)qinfo[0].qid;
                atomicAdd(&u[0], a_hits);
                atomicAdd(&u[1], (unsigned long long)a_nseq);
                atomicAdd(&u[2], (unsigned long long)a_ext);
                atomicAdd(&u[3], (unsigned long long)a_saved);
                atomicAdd(&u[4], (unsigned long long)a_passed);
            }
        } else {
#pragma unroll
            for (int t = 0; t < kQuerySlots; ++t) {
                const uint32_t jj = lane + 64u * (uint32_t)t;
                if (jj < g.nq && m_nseq[t]) {
                    unsigned long long* u = p.ustats + 5 * (size_t)qinfo[jj].qid;
                    atomicAdd(&u[0], m_hits[t]);
                    atomicAdd(&u[1], (unsigned long long)m_nseq[t]);
                    atomicAdd(&u[2], (unsigned long long)m_ext[t]);
                    atomicAdd(&u[3], (unsigned long long)m_saved[t]);
                    atomicAdd(&u[4], (unsigned long long)m_passed[t]);
                }
            }
        }
        zero_ustats();
    };
    zero_ustats();

    uint32_t pass = 0;               // flat: items in order, then done
    uint32_t c_next = 0, c_end = 0;  // flat: this wave's claimed subjects (wave-uniform)
    for (;;) {
        // Flat: every wavefront walks the items in order on its own and the
        // workgroup meets only where the group, and with it the LDS staging,
        // changes -- consecutive items of one group (cost bins of its
        // subjects) run without a barrier.  Otherwise the workgroup claims
        // items one at a time from the launch's counter.
        uint32_t item;
        if (p.flat) {
            item = min(pass, p.nitems);
        } else {
            __syncthreads();
            if (tid == 0) {
                ctl[0] = atomicAdd(p.work_counter, 1u);
                ctl[1] = 0;
            }
            __syncthreads();
            item = ctl[0];
        }
        ++pass;
        c_next = c_end = 0;
        // flat: this wave is done with its group when the next item belongs
        // to another (each wave reads qinfo before the barrier of the change)
        if (p.flat && cur_group != 0xFFFFFFFFu && (item >= p.nitems || p.items[item].group != cur_group))
            flush_ustats();
        if (item >= p.nitems) {
            // largest record count any wave produced: sizes a rescan after
            // a segment overflow exactly
            if (lane == 0 && seg_fill > p.seg_cap) atomicMax(p.max_fill, seg_fill);
            if (timed && lane == 0) {
                unsigned long long* ph = p.phase + 8 * (size_t)(p.seg_wave_base + blockIdx.x * nwaves + wave);
                ph[0] = t_begin;
                ph[1] = __builtin_amdgcn_s_memrealtime();
                ph[2] = t_stage;
                ph[3] = t_group;
                ph[4] = n_staged;
                ph[5] = b_staged;
                ph[6] = t_flush;
                ph[7] = n_flushed;
            }
            break;
        }
        const WorkItem w = p.items[item];
        // subjects a wavefront claims at a time from this item (flat)
        const uint32_t chunk = w.chunk ? w.chunk : p.flat;
        const uint64_t t_grp0 = timed ? __builtin_amdgcn_s_memrealtime() : 0;
        if (w.group != cur_group) {
            // every wave is done with the previous group's LDS data
            if (p.flat) __syncthreads();
            g = p.groups[w.group];
            if (TAB_LDS)
                for (uint32_t i = tid; i < g.table_words; i += blockDim.x) tab_lds[i] = p.tables[g.table_word + i];
            for (uint32_t i = tid; i < g.qres_len; i += blockDim.x) Q[i] = p.qres[g.qres_off + i];
            if (MULTI)
                for (uint32_t i = tid; i < g.qres_len; i += blockDim.x) qmap[i] = p.qmap[g.qres_off + i];
            for (uint32_t i = tid; i < g.nq; i += blockDim.x) qinfo[i] = p.qinfo[g.qinfo_first + i];
            cur_group = w.group;
            __syncthreads();
            if (timed) t_group += __builtin_amdgcn_s_memrealtime() - t_grp0;
        }
        const uint32_t* table = TAB_LDS ? tab_lds : p.tables + g.table_word;
        const uint2* pvr = reinterpret_cast<const uint2*>(table);
        // start offsets are u16 (padded to 4 B) for tables of < 65,536
        // entries, which every LDS-resident table is, else u32
        const bool narrow_starts = TAB_LDS || (g.flags & kGroupNarrowStarts) != 0;
        const uint32_t* starts32 = table + 2 * p.ngroups_pv;
        const uint16_t* starts16 = reinterpret_cast<const uint16_t*>(starts32);
        const uint16_t* entries = narrow_starts ? starts16 + ((g.npresent + 2u) & ~1u)
                                                : reinterpret_cast<const uint16_t*>(starts32 + g.npresent + 1);
        // a table in HBM is looked up through its start offsets by word
        // (one load for a word's range instead of presence word, rank and
        // starts in turn); words at or past wlast read an empty range
        const uint32_t* direct = p.tables + g.direct_word;
        const uint32_t wlast = p.ngroups_pv * 32u;
        const QueryInfo q0 = qinfo[0];
        // flat: the descriptors of the claimed chunk, lane j holding those of
        // its j-th subject -- loaded with the claim, so a subject's residue
        // loads wait for one memory round trip, not two
        uint32_t c_first = 0, d_sid = 0, d_len = 0;
        uint64_t d_off = 0;

        for (;;) {
            uint32_t k = 0;
            const uint64_t t_st0 = timed ? __builtin_amdgcn_s_memrealtime() : 0;
            if (p.flat) {
                // a chunk of the item's subjects from its device-wide counter:
                // every subject goes to exactly one wave, in processing order
                if (c_next == c_end) {
                    uint32_t c = 0;
                    if (lane == 0) c = atomicAdd(&p.flat_counters[item], chunk);
                    c_next = (uint32_t)__builtin_amdgcn_readlane((int)c, 0);
                    c_end = c_next + chunk;
                    c_first = c_next;
                    // descriptors are stored in processing order: a chunk
                    // reads consecutive entries (chunk <= 16 lanes)
                    const uint32_t kk = c_next + lane;
                    if (lane < chunk && kk < w.count) {
                        const uint32_t o = w.first + kk;
                        d_sid = p.order[o];
                        d_off = p.poff[o];
                        d_len = p.plen[o];
                    }
                }
                k = c_next++;
            } else {
                if (lane == 0) k = atomicAdd(&ctl[1], 1u);
                k = (uint32_t)__builtin_amdgcn_readlane((int)k, 0);
            }
            if (k >= w.count) break;
            uint32_t sid;
            uint64_t soff;
            int32_t slen;
            if (p.flat) {
                const int j = (int)(k - c_first);  // wave-uniform
                sid = (uint32_t)__builtin_amdgcn_readlane((int)d_sid, j);
                soff = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)d_off, j) |
                       ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(d_off >> 32), j) << 32);
                slen = __builtin_amdgcn_readlane((int)d_len, j);
            } else {
                const uint32_t opos = w.first + k;  // position in the processing order
                sid = p.order[opos];
                soff = p.poff[opos];
                slen = (int32_t)p.plen[opos];
            }
            const uint32_t slot = sid - p.first;
            const uint8_t* __restrict__ S = p.residues + soff;
            const int32_t send = (slen + 15) & ~15;
            const bool resident = send <= (int32_t)p.sbuf;

            // ---- the pack: this subject and, while they fit the subject
            // buffer beside it, the next ones of the claimed chunk ----
            // The wave scans them as one stream: subject t sits at buffer
            // offset V_t (16-byte aligned, at least length + window past the
            // previous subject's start), and a hit of subject t at s is a
            // hit at V_t + s.  V_t - V_{t-1} >= length + window is NCBI's
            // running diagonal offset between two subjects
            // (Blast_ExtendWordExit), so every slot an earlier subject left
            // reads as stale, exactly as after the per-subject reset.  The
            // ring, the extension queue and the deferred hits carry across
            // the subjects of a pack; its extensions run in shared flushes.
            int32_t V1 = INT32_MAX, V2 = INT32_MAX, V3 = INT32_MAX;  // offsets (V_0 = 0)
            int32_t N1 = 0, N2 = 0, N3 = 0;                         // lengths
            uint32_t Z1 = 0, Z2 = 0, Z3 = 0;                        // slots
            uint64_t O1 = 0, O2 = 0, O3 = 0;                        // arena offsets
            int32_t npack = 1, vend = slen;                         // subjects, stream end
            if (p.flat && resident) {
                while (npack < kPack && c_next < c_end && c_next < w.count) {
                    const int j = (int)(c_next - c_first);  // wave-uniform
                    const int32_t ln = __builtin_amdgcn_readlane((int)d_len, j);
                    const int32_t v = (vend + window + 15) & ~15;
                    if (v + ((ln + 15) & ~15) > (int32_t)p.sbuf) break;
                    const uint32_t z = (uint32_t)__builtin_amdgcn_readlane((int)d_sid, j) - p.first;
                    const uint64_t o =
                        (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)d_off, j) |
                        ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(d_off >> 32), j) << 32);
                    if (npack == 1) {
                        V1 = v; N1 = ln; Z1 = z; O1 = o;
                    } else if (npack == 2) {
                        V2 = v; N2 = ln; Z2 = z; O2 = o;
                    } else {
                        V3 = v; N3 = ln; Z3 = z; O3 = o;
                    }
                    vend = v + ln;
                    ++npack;
                    ++c_next;
                }
            }
            // the pack subject of stream position v
            auto sub_of = [&](int32_t v) { return (int32_t)(v >= V1) + (int32_t)(v >= V2) + (int32_t)(v >= V3); };

            // ---- stage the pack (or a long subject's first window) into LDS ----
            auto fill = [&](int32_t from, int32_t to) {
                for (int32_t x = from + 16 * (int32_t)lane; x < to; x += 1024)
                    *reinterpret_cast<uint4*>(sbuf + (x & smask)) = *reinterpret_cast<const uint4*>(S + x);
            };
            if (NARROW && doff + vend + window > (int32_t)kVal) {
                clear_diag();
                doff = window;
            }
            int32_t whi = min(send, (int32_t)p.sbuf);
            uint64_t staged = (uint64_t)whi;
            if (npack == 1) {
                fill(0, whi);
            } else {
                // every subject's first 1 KB is loaded before any is stored,
                // so the pack's loads are in flight together
                const uint8_t* __restrict__ S1 = p.residues + O1;
                const uint8_t* __restrict__ S2 = p.residues + O2;
                const uint8_t* __restrict__ S3 = p.residues + O3;
                const int32_t e1 = (N1 + 15) & ~15, e2 = (N2 + 15) & ~15, e3 = (N3 + 15) & ~15;
                const int32_t x = 16 * (int32_t)lane;
                uint4 r0{}, r1{}, r2{}, r3{};
                if (x < send) r0 = *reinterpret_cast<const uint4*>(S + x);
                if (x < e1) r1 = *reinterpret_cast<const uint4*>(S1 + x);
                if (x < e2) r2 = *reinterpret_cast<const uint4*>(S2 + x);
                if (x < e3) r3 = *reinterpret_cast<const uint4*>(S3 + x);
                if (x < send) *reinterpret_cast<uint4*>(sbuf + x) = r0;
                if (x < e1) *reinterpret_cast<uint4*>(sbuf + V1 + x) = r1;
                if (x < e2) *reinterpret_cast<uint4*>(sbuf + V2 + x) = r2;
                if (x < e3) *reinterpret_cast<uint4*>(sbuf + V3 + x) = r3;
                for (int32_t y = x + 1024; y < send; y += 1024)
                    *reinterpret_cast<uint4*>(sbuf + y) = *reinterpret_cast<const uint4*>(S + y);
                for (int32_t y = x + 1024; y < e1; y += 1024)
                    *reinterpret_cast<uint4*>(sbuf + V1 + y) = *reinterpret_cast<const uint4*>(S1 + y);
                for (int32_t y = x + 1024; y < e2; y += 1024)
                    *reinterpret_cast<uint4*>(sbuf + V2 + y) = *reinterpret_cast<const uint4*>(S2 + y);
                for (int32_t y = x + 1024; y < e3; y += 1024)
                    *reinterpret_cast<uint4*>(sbuf + V3 + y) = *reinterpret_cast<const uint4*>(S3 + y);
                staged += (uint64_t)(e1 + e2 + e3);
                // the gaps between the pack's subjects hold code 31, so no
                // word that reaches into a gap is present in the table
                auto gap = [&](int32_t from, int32_t to) {
                    for (int32_t y = from + (int32_t)lane; y < to; y += 64) sbuf[y] = (uint8_t)31;
                };
                gap(slen, V1);
                if (npack > 2) gap(V1 + N1, V2);
                if (npack > 3) gap(V2 + N2, V3);
            }
            wave_sync();
            if (timed) {
                t_stage += __builtin_amdgcn_s_memrealtime() - t_st0;
                n_staged += (uint64_t)npack;
                b_staged += staged;
            }

            // single-query word hits, accumulated for pack subject hsub and
            // stored to pk[] when the scan passes into the next subject
            uint32_t hits_cur = 0;
            int32_t hsub = 0;
            const uint32_t rec_start = seg_fill;  // the pack's run of records
            uint32_t head = 0, count = 0;                 // hit ring (wave-uniform)
            uint32_t ntasks = 0, ndef = 0;                // queued extensions, deferred hits
            uint32_t rep = 0, rep_total = 0;              // deferred hits being re-applied
            uint32_t carry = 0;                           // ring-front entries left unapplied by the last batch
            int32_t qbase = 0;

            // Applies n (<= 64) hits to the diagonal table; lane i holds the
            // i-th in scan order.  Hits on one slot are applied in lane order
            // (the lowest lane wins a hashed claim each round).  A hit that
            // completes a pair marks its slot pending and queues the
            // extension; a hit that finds its slot pending is deferred --
            // appended, in order, to the deferred queue, which is applied
            // again right after the next flush.  So every slot still sees its
            // hits strictly in scan order, and a flush is needed only when a
            // queue is full or the subject ends.  The caller guarantees
            // ntasks + n <= kTaskCap and ndef + n <= kDefCap.
            //
            // hook(r), r = 0, 1, runs right after round r's state reads are
            // issued (each once, in order; after the rounds when the batch is
            // done early or has fewer rounds): the next positions' lookups go
            // there, so their LDS latency overlaps the claim rounds'.
            //
            // `rounds` (a compile-time count) bounds the claim rounds: a ring
            // batch runs two, and the hits still unapplied -- each lost to a
            // lower lane on its slot or to a hash collision in both rounds --
            // go back to the front of the ring, ahead of every newer hit, and
            // ride in the next batch (lanes below fresh_from of that batch).
            // A slot's hits keep their order: the lanes left over are later
            // than every applied hit of their slot and earlier than every hit
            // behind them in the ring.  Most batches then need two rounds
            // instead of three or four (chains of hits on one diagonal from
            // overlapping words are common).  Re-applied deferred hits run
            // to completion (rounds = 0).
            auto apply = [&](uint32_t e, uint32_t n, uint32_t fresh_from, auto rounds, auto&& hook) {
                constexpr int kRounds = decltype(rounds)::value;
                int hooked = 0;
                bool active = lane < n;
                const bool fresh = lane >= fresh_from;
                const int32_t c = (int32_t)(e >> 16);
                const int32_t s = qbase + (int32_t)(e & 0xFFFFu);
                uint32_t j = 0, dbase = q0.dbase, dmask = q0.dmask;
                int32_t qstart = (int32_t)q0.qstart;
                if (MULTI) {
                    j = qmap[c];
                    const QueryInfo qi = qinfo[j];
                    dbase = qi.dbase;
                    dmask = qi.dmask;
                    qstart = (int32_t)qi.qstart;
                    if (active && fresh) atomicAdd(&qc[__umul24((uint32_t)sub_of(s), p.max_nq) + j].hits, 1u);
                }
                const uint32_t dslot = dbase + ((uint32_t)((c - qstart) - s) & dmask);
                // The first round hashes the slot by its low bits; later rounds
                // by a multiplicative hash, so lanes that lost only to a hash
                // collision rarely collide again.  Lanes on one slot share
                // every round's hash, so the lowest still goes first.
                const uint32_t cmask = p.claim_slots - 1;
                uint32_t h = dslot & cmask;
                const uint32_t h2 = (__umul24(dslot, 0x9E3779u) >> 16) & cmask;  // 24-bit multiply: full rate
                bool deferred = false;
                const int32_t sd = s + doff;
                // (n >= 1: round 0 always has an active lane)
                for (int round = 0; (kRounds == 0 || round < kRounds) && (round == 0 || __ballot(active) != 0);
                     ++round) {
                    if (round > 0) h = h2;
                    if (active) atomicMin(&claim[h], lane);
                    wave_sync();
                    // slot state and claim are read by every lane, together;
                    // the state machine below is evaluated without branches
                    const uint32_t cl = claim[h];
                    const uint32_t st = diag[dslot];
                    if (kRounds > 0 && round < kRounds) {
                        hook(round);
                        hooked = round + 1;
                    }
                    wave_sync();
                    const bool win = active && cl == lane;
                    const int32_t v = (int32_t)(st & kVal);
                    const int32_t diff = sd - v;  // s - last_hit
                    const bool live = win && (st & kPend) == 0;
                    const bool flagged = (st & kFlag) != 0;
                    // flagged: the hit restarts the diagonal once it is past
                    // the extension's end; unflagged: a hit beyond the window
                    // restarts it, one at least a word past the last completes
                    // a pair
                    const bool restart = diff >= (flagged ? 0 : window);
                    const bool task = live && !flagged && diff < window && diff >= WORD;
                    const bool write = live && (restart || task);
                    const uint32_t ns = task ? (st | kPend) : (uint32_t)sd;
                    deferred = deferred || (win && !live);
                    active = active && !win;
                    if (win) claim[h] = 0xFFFFFFFFu;
                    const unsigned long long tb = __ballot(task);
                    if (tb) {
                        if (task) {
                            const uint32_t idx = ntasks + rank_below(tb);
                            t_sc[idx] = dslot | ((uint32_t)c << 16);
                            t_s[idx] = (uint16_t)(s - qbase);
                        }
                        ntasks += (uint32_t)__popcll(tb);
                    }
                    if (write) diag[dslot] = (SlotT)ns;
                    wave_sync();
                }
                if (kRounds > 0)
                    for (int r = hooked; r < 2; ++r) hook(r);
                if (kRounds > 0) {
                    // lanes still unapplied go back to the front of the ring,
                    // in order: they precede every hit still in it
                    const unsigned long long lb = __ballot(active);
                    const uint32_t nl = (uint32_t)__popcll(lb);
                    if (active) ring[(head - nl + rank_below(lb)) & qmask] = e;
                    head = (head - nl) & qmask;
                    count += nl;
                    carry = nl;
                    wave_sync();
                }
                const unsigned long long db = __ballot(deferred);
                if (deferred) dq[ndef + rank_below(db)] = e;
                ndef += (uint32_t)__popcll(db);
                wave_sync();
            };

            // Runs the first `take` queued extensions and writes their
            // outcome to the diagonal table; the others move to the front of
            // the queue, in order (their slots stay pending, and hits on them
            // stay deferred).  A few tasks run on kSeg-lane segments
            // (64 / kSeg per pass); more run one per lane.
            auto flush = [&](uint32_t take) {
                const uint32_t nt = take;
                const uint64_t t_f0 = timed ? __builtin_amdgcn_s_memrealtime() : 0;
                wave_sync();
                constexpr uint32_t kPer = 64u / kSeg;
                const bool coop = resident && nt <= kCoopMax;
                const uint32_t npass = coop ? (nt + kPer - 1) / kPer : (nt + 63) / 64;
                for (uint32_t pass = 0; pass < npass; ++pass) {
                    const uint32_t t = coop ? pass * kPer + lane / kSeg : pass * 64 + lane;
                    const bool tvalid = t < nt;
                    const bool mine = coop ? (tvalid && (lane & (kSeg - 1u)) == 0) : tvalid;
                    const uint32_t ti = min(t, kTaskCap - 1);
                    const uint32_t sc = tvalid ? t_sc[ti] : 0u;  // slot 0 / position 0 for idle lanes
                    const uint32_t tslot = sc & 0xFFFFu;
                    int32_t tc = (int32_t)(sc >> 16);
                    const uint32_t tj = MULTI ? (uint32_t)qmap[tc] : 0u;
                    const int32_t tv_s = qbase + (int32_t)t_s[ti];  // stream position of the second hit
                    // its pack subject: buffer offset, length, slot; the
                    // extension runs in that subject's own coordinates
                    const int32_t tp = sub_of(tv_s);
                    const int32_t tv = tp == 0 ? 0 : (tp == 1 ? V1 : (tp == 2 ? V2 : V3));
                    const int32_t tn = tp == 0 ? slen : (tp == 1 ? N1 : (tp == 2 ? N2 : N3));
                    const uint32_t tz = tp == 0 ? slot : (tp == 1 ? Z1 : (tp == 2 ? Z2 : Z3));
                    const int32_t ts = tv_s - tv;
                    // left boundary: the first hit (kept in the slot) + word
                    const int32_t tsl = (int32_t)(diag[tslot] & kVal) - doff - tv + WORD;
                    const QueryInfo qi = MULTI ? qinfo[tj] : q0;
                    TwoHit x{};
                    if (coop) {
                        x = coop_extend_two_hit<WORD, kSeg>(M, Q, p.qres_cap, sbuf, smask, (uint32_t)tv, tn, tvalid,
                                                            (int32_t)qi.qstart, (int32_t)qi.qlen, tsl, ts, tc,
                                                            qi.xdrop, lane);
                    } else if (mine) {
                        if (resident)
                            x = extend_resident<WORD>(M, Q, (int32_t)qi.qstart, (int32_t)qi.qlen, sbuf, smask,
                                                      (uint32_t)tv, tn, tsl, ts, tc, qi.xdrop);
                        else  // a long subject: always alone in its pack
                            x = extend_two_hit<WORD, false>(M, Q, p.qres_cap, (int32_t)qi.qstart, (int32_t)qi.qlen,
                                                            sbuf, smask, S, whi - (int32_t)p.sbuf, whi, slen, tsl, ts,
                                                            tc, qi.xdrop);
                    }
                    bool rec = false;
                    if (mine) {
                        rec = x.score >= qi.cutoff;
                        diag[tslot] = (SlotT)(x.right_extend ? ((uint32_t)(x.s_last + tv - (WORD - 1) + doff) | kFlag)
                                                             : (uint32_t)(tv_s + doff));
                        if (MULTI) {
                            QCount& cq = qc[__umul24((uint32_t)tp, p.max_nq) + tj];
                            atomicAdd(&cq.ext, 1u);
                            if (rec) atomicAdd(&cq.saved, 1u);
                        }
                        tc -= (int32_t)qi.qstart;  // query-local offset of the second hit
                    }
                    const unsigned long long rb = __ballot(rec);
                    if (!MULTI) {
                        // per pack subject: extensions run and records kept
                        if (npack == 1) {
                            const unsigned long long mb = __ballot(mine);
                            if (lane == 0 && mb) {
                                pk[0].ext += (uint32_t)__popcll(mb);
                                pk[0].saved += (uint32_t)__popcll(rb);
                            }
                        } else {
                            for (int u = 0; u < npack; ++u) {
                                const unsigned long long mi = __ballot(mine && tp == u);
                                if (lane == 0 && mi) {
                                    pk[u].ext += (uint32_t)__popcll(mi);
                                    pk[u].saved += (uint32_t)__popcll(rb & mi);
                                }
                            }
                        }
                    }
                    if (rb) {
                        const uint32_t at = seg_fill + rank_below(rb);
                        if (rec) {
                            if (at < p.seg_cap) {
                                SegRecord r;
                                r.h.hsp_q = x.hsp_q;
                                r.h.hsp_s = x.hsp_s;
                                r.h.query_offset = tc;
                                r.h.subject_offset = ts;
                                r.h.hsp_len = x.len;
                                r.h.score = x.score;
                                r.j = tj;
                                r.slot = tz;
                                p.seg[seg_base + at] = r;
                            } else {
                                atomicOr(p.overflow, 1u);
                            }
                        }
                        seg_fill += (uint32_t)__popcll(rb);
                    }
                }
                // the rest (at most 64: take = 64 < ntasks <= kTaskCap) to
                // the front; the source [nt, ntasks) lies past the
                // destination [0, ntasks - nt) since nt >= ntasks - nt
                const uint32_t rest = ntasks - nt;
                if (rest) {
                    uint32_t a = 0, b = 0;
                    if (lane < rest) {
                        a = t_sc[nt + lane];
                        b = t_s[nt + lane];
                    }
                    wave_sync();
                    if (lane < rest) {
                        t_sc[lane] = a;
                        t_s[lane] = (uint16_t)b;
                    }
                }
                ntasks = rest;
                wave_sync();
                if (timed) {
                    t_flush += __builtin_amdgcn_s_memrealtime() - t_f0;
                    n_flushed += nt;
                }
            };

            // Positions are consumed 64 at a time into the ring.  Each pass
            // of the loop does one of: re-apply up to 64 deferred hits (right
            // after a flush, before anything else); flush (a queue cannot
            // take the next batch, or everything must be applied: subject
            // end, 16-bit offset rebase); apply the next ring batch (a full
            // batch is ready, or the ring must be emptied); fill the ring
            // from the next 64 positions.
            const int32_t last = vend - WORD;  // the pack's last word position
            int32_t pos = 0;
            bool empty_ring = false, full_drain = false;

            // The lookup of positions pos .. pos + 63 (one per lane) in four
            // stages, each waiting for the previous one's reads: the word's
            // residues; its presence bits and rank base (a table in HBM: its
            // cell's range, read by word); its cell's start offsets (a table
            // in HBM: its first eight entries); the hits into the ring.
            struct Look {
                uint32_t b[WORD];
                uint32_t wd, start, end;
                uint2 gv;
                uint32_t pe[8];  // a table in HBM: the cell's first entries
            };
            auto look_a = [&](Look& L) {
                const int32_t s = pos + (int32_t)lane;
#pragma unroll
                for (int i = 0; i < WORD; ++i) L.b[i] = sbuf[(s + i) & smask];
            };
            auto look_b = [&](Look& L) {
                uint32_t wd = L.b[0];
#pragma unroll
                for (int i = 1; i < WORD; ++i) wd = (wd << 5) | L.b[i];
                // lanes past the subject's end read buffer bytes that may
                // never have been written (any value): keep their word in
                // range so its presence-word read stays inside the table
                wd &= (1u << (5 * WORD)) - 1u;
                // a word holding a pack's gap byte (code 31) is absent: as a
                // later residue its presence bit is never set (no table word
                // holds code 31); as the first, the clamp below lands on the
                // presence word of (27, 31, *) for 3-letter words, also
                // empty; 2-letter words would land on (27, *), so they are
                // sent to (27, 31)
                if (WORD == 2 && L.b[0] >= 28u) wd = (27u << 5) | 31u;
                L.wd = wd;
                if (TAB_LDS) {
                    L.gv = pvr[min(wd >> 5, p.ngroups_pv - 1u)];
                } else {
                    const uint32_t wi = min(wd, wlast);
                    L.start = direct[wi];
                    L.end = direct[wi + 1];
                }
            };
            auto look_c = [&](Look& L) {
                if (!TAB_LDS) {
                    // the range was read by word: the cell's first eight
                    // entries go out now, so their memory round trip
                    // overlaps the rest of the claim rounds (most cells hold
                    // no more).  Loads past the cell stay inside the table
                    // buffer, as in the ring append.
#pragma unroll
                    for (uint32_t k = 0; k < 8; ++k) L.pe[k] = entries[L.start + k];
                    return;
                }
                const uint32_t bit = 1u << (L.wd & 31u);
                // rank of the word among present words; an absent word reads
                // the neighbouring (valid) start entries and is masked out
                const uint32_t r =
                    min(L.gv.y + (uint32_t)__popc(L.gv.x & (bit - 1u)), g.npresent ? g.npresent - 1u : 0u);
                if (narrow_starts) {
                    L.start = starts16[r];
                    L.end = starts16[r + 1];
                } else {
                    L.start = starts32[r];
                    L.end = starts32[r + 1];
                }
            };
            // appends the hits of as many leading positions as the ring holds
            auto look_d = [&](Look& L) {
                const int32_t s = pos + (int32_t)lane;
                const bool present = (!TAB_LDS || (L.gv.x & (1u << (L.wd & 31u))) != 0) && s <= last;
                const uint32_t cnt = present ? L.end - L.start : 0u;
                const uint32_t incl = wave_inclusive_sum(cnt);
                const bool fits = incl <= qcap - count;
                const uint32_t nfit = (uint32_t)__popcll(__ballot(fits));
                if (nfit == 0) {
                    if (count == 0) {  // a cell larger than the ring: host sizing error
                        if (lane == 0) atomicOr(p.overflow, 2u);
                        pos = last + 1;
                    } else {
                        empty_ring = true;
                    }
                    return;
                }
                if (fits && cnt) {
                    const uint32_t at = head + count + (incl - cnt);
                    const uint32_t sv = (uint32_t)(s - qbase);
                    if (TAB_LDS) {
                        for (uint32_t t = 0; t < cnt; ++t)
                            ring[(at + t) & qmask] = ((uint32_t)entries[L.start + t] << 16) | sv;
                    } else {
                        // a table in HBM: the first eight entries were loaded
                        // during the claim rounds; the rest go out eight
                        // independent loads per step, so a cell costs one
                        // memory round trip per eight entries instead of one
                        // per entry.  Loads past the cell stay inside the
                        // table buffer (the next group's table, or the start
                        // offsets by word after the last one).
#pragma unroll
                        for (uint32_t k = 0; k < 8; ++k)
                            if (k < cnt) ring[(at + k) & qmask] = (L.pe[k] << 16) | sv;
                        for (uint32_t t = 8; t < cnt; t += 8) {
                            uint32_t v[8];
#pragma unroll
                            for (uint32_t k = 0; k < 8; ++k) v[k] = entries[L.start + t + k];
#pragma unroll
                            for (uint32_t k = 0; k < 8; ++k)
                                if (t + k < cnt) ring[(at + t + k) & qmask] = (v[k] << 16) | sv;
                        }
                    }
                }
                // nfit is wave-uniform: a lane read, not an LDS permute
                const uint32_t added = (uint32_t)__builtin_amdgcn_readlane((int)incl, (int)nfit - 1);
                count += added;
                if (!MULTI) {
                    // the hits of lanes past a subject boundary belong to the
                    // next subject(s) of the pack
                    uint32_t taken = 0;
                    for (;;) {
                        const int32_t vb = hsub == 0 ? V1 : (hsub == 1 ? V2 : (hsub == 2 ? V3 : INT32_MAX));
                        if (pos + (int32_t)nfit <= vb) break;
                        const int32_t lb = max(vb - pos, 0);  // lanes before the boundary
                        const uint32_t h = lb > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)incl, lb - 1) : taken;
                        hits_cur += h - taken;
                        if (lane == 0) pk[hsub].hits = hits_cur;
                        taken = h;
                        hits_cur = 0;
                        ++hsub;
                    }
                    hits_cur += added - taken;
                }
                pos += (int32_t)nfit;
                wave_sync();
            };
            for (;;) {
                uint32_t n, e;
                if (rep < rep_total) {
                    n = min(64u, rep_total - rep);
                    e = lane < n ? dq[rep + lane] : 0u;
                    rep += n;
                    apply(e, n, 64u, std::integral_constant<int, 0>{}, [](int) {});
                } else {
                    const bool at_end = pos > last;
                    const bool drain_all = at_end || full_drain;
                    n = min(count, 64u);
                    const bool batch = count >= 64 || ((empty_ring || drain_all) && count > 0);
                    if (batch ? (ntasks + n > kTaskCap || ndef + n > kDefCap) : (drain_all && ntasks > 0)) {
                        // a queue-full flush runs one full pass of 64 and
                        // keeps the rest queued (a second, partly idle pass
                        // costs about as much as a full one), unless the
                        // kept tasks and one task per deferred hit might not
                        // fit the queue when the deferred hits are re-applied
                        flush(!drain_all && ntasks > 64u && ntasks - 64u + ndef <= kTaskCap ? 64u : ntasks);
                        rep = 0;
                        rep_total = ndef;
                        ndef = 0;
                        continue;
                    }
                    if (!batch) {
                        empty_ring = full_drain = false;
                        if (at_end) break;
                        if (pos + 128 > whi && whi < send) {
                            const uint64_t t_w0 = timed ? __builtin_amdgcn_s_memrealtime() : 0;
                            const int32_t nhi = min(send, whi + (int32_t)(p.sbuf >> 1));
                            fill(whi, nhi);
                            b_staged += timed ? (uint64_t)(nhi - whi) : 0u;
                            whi = nhi;
                            wave_sync();
                            if (timed) t_stage += __builtin_amdgcn_s_memrealtime() - t_w0;
                        }
                        if (count == 0 && ndef == 0 && ntasks == 0) {
                            qbase = pos;
                        } else if (pos + 64 - qbase >= 65536) {
                            empty_ring = full_drain = true;
                            continue;
                        }
                        Look L;
                        look_a(L);
                        look_b(L);
                        look_c(L);
                        look_d(L);
                        continue;
                    }
                    e = lane < n ? ring[(head + lane) & qmask] : 0u;
                    head = (head + n) & qmask;
                    count -= n;
                    // a full batch, and the next 64 positions can be looked up
                    // without a window refill or a rebase: their lookups run
                    // interleaved with the batch's claim rounds.  Only for a
                    // table in LDS: an HBM table's lookups would stall the
                    // rounds (measured 6 % slower on batched scans)
                    const bool combo = TAB_LDS && n == 64 && !drain_all && !empty_ring && pos <= last &&
                                       (pos + 128 <= whi || whi >= send) && pos + 64 - qbase < 65536;
                    if (combo) {
                        Look L;
                        look_a(L);
                        apply(e, n, carry, std::integral_constant<int, 2>{}, [&](int stage) {
                            if (stage == 0) look_b(L); else look_c(L);
                        });
                        look_d(L);
                    } else {
                        apply(e, n, carry, std::integral_constant<int, 2>{}, [](int) {});
                    }
                }
            }

            // ---- the pack's outputs ----
            // Counts and run lengths were zero-filled before the launch, so
            // only subjects with records write them.  The per-subject
            // counters go straight to pinned host memory at the subject's
            // slot: 12-byte posted writes over the host link while the scan
            // runs, instead of a permute and a copy after it.  The pack's
            // records form one run, filed under its first subject; the gather
            // sorts it by subject.
            const uint32_t gslot = w.group * p.count + slot;
            const bool has_run = seg_fill > rec_start;
            if (!MULTI) {
                if (lane == 0) pk[hsub].hits = hits_cur;
                wave_sync();
                for (int u = 0; u < npack; ++u) {
                    const PackSub c = pk[u];
                    const uint32_t z = u == 0 ? slot : (u == 1 ? Z1 : (u == 2 ? Z2 : Z3));
                    if (lane == 0) {
                        if (c.saved) p.cnt[(size_t)q0.qid * p.count + z] = c.saved;
                        if (p.stats) {
                            int32_t* st = p.stats + 3 * ((size_t)q0.qid * p.count + z);
                            st[0] = (int32_t)c.hits;
                            st[1] = (int32_t)c.saved;
                            st[2] = (int32_t)c.ext;
                        }
                    }
                    if (c.hits > 0) {  // Blast_UngappedStatsUpdate skips subjects without hits
                        a_hits += c.hits;
                        a_nseq += 1;
                        a_ext += c.ext;
                        a_saved += c.saved;
                        a_passed += c.saved > 0;
                    }
                }
                wave_sync();
                if (lane < (uint32_t)npack) pk[lane] = PackSub{};
                if (lane == 0 && has_run) {
                    p.run_loc[gslot] = seg_base + rec_start;
                    p.run_len[gslot] = seg_fill - rec_start;
                }
                wave_sync();
            } else {
                wave_sync();
                for (int u = 0; u < npack; ++u) {
                    const uint32_t z = u == 0 ? slot : (u == 1 ? Z1 : (u == 2 ? Z2 : Z3));
                    QCount* qe = qc + u * p.max_nq;
#pragma unroll
                    for (int t = 0; t < kQuerySlots; ++t) {
                        const uint32_t jj = lane + 64u * (uint32_t)t;
                        if (jj < g.nq) {
                            QCount& cq = qe[jj];
                            const size_t qrow = (size_t)qinfo[jj].qid * p.count;
                            if (cq.saved) p.cnt[qrow + z] = cq.saved;
                            if (p.stats) {
                                int32_t* st = p.stats + 3 * (qrow + z);
                                st[0] = (int32_t)cq.hits;
                                st[1] = (int32_t)cq.saved;
                                st[2] = (int32_t)cq.ext;
                            }
                            if (cq.hits > 0) {
                                m_hits[t] += cq.hits;
                                m_nseq[t] += 1;
                                m_ext[t] += cq.ext;
                                m_saved[t] += cq.saved;
                                m_passed[t] += cq.saved > 0;
                            }
                            cq.hits = cq.ext = cq.saved = 0;
                        }
                    }
                }
                if (lane == 0 && has_run) {
                    p.run_loc[gslot] = seg_base + rec_start;
                    p.run_len[gslot] = seg_fill - rec_start;
                }
                wave_sync();
            }
            // Blast_ExtendWordExit: a fresh diagonal range for the next
            // pack -- NCBI's running offset, advanced by the pack's stream
            // length + window, so every slot a previous subject left reads as
            // a hit at least a window back (or a flag behind the next hit).
            // The table is cleared only when the offset would outgrow the
            // slot's value bits: narrow slots check at the next subject's
            // start (its length decides), wide slots here.
            if (!NARROW && doff >= kOffsetLimit) {
                clear_diag();
                doff = window;
                wave_sync();
            } else {
                doff += vend + window;
            }
        }

        // claimed items: the next item may belong to another group
        if (!p.flat) flush_ustats();
    }
}

// ---- compaction ----------------------------------------------------------------

constexpr uint32_t kCompactThreads = 256;
constexpr uint32_t kPer = kTile / kCompactThreads;  // 4 counts per thread

__device__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* sh, uint32_t* total) {
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v);
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

__global__ __launch_bounds__(kCompactThreads) void k_tile_sums(CompactParams p) {
    __shared__ uint32_t sh[kCompactThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < kPer; ++k)
        if (base + k < p.n) sum += p.cnt[base + k];
    uint32_t tot;
    block_exclusive_sum(sum, sh, &tot);
    if (threadIdx.x == 0) p.tile_sum[blockIdx.x] = tot;
}

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
    if (threadIdx.x == 0) *p.total = carry;
}

// counts -> exclusive offsets, in place
__global__ __launch_bounds__(kCompactThreads) void k_offsets(CompactParams p) {
    __shared__ uint32_t sh[kCompactThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t c[kPer], sum = 0;
    for (uint32_t k = 0; k < kPer; ++k) {
        c[k] = base + k < p.n ? p.cnt[base + k] : 0u;
        sum += c[k];
    }
    uint32_t tot;
    uint32_t off = p.tile_sum[blockIdx.x] + block_exclusive_sum(sum, sh, &tot);
    for (uint32_t k = 0; k < kPer; ++k)
        if (base + k < p.n) {
            p.cnt[base + k] = off;
            off += c[k];
        }
}

// Record order within a (group, subject) run: query, subject offset, query
// offset.  Keys are unique (a second hit gives at most one record per query).
__device__ __forceinline__ bool rec_before(const SegRecord& a, const SegRecord& b) {
    if (a.slot != b.slot) return a.slot < b.slot;
    return a.j < b.j || (a.j == b.j && (a.h.subject_offset < b.h.subject_offset ||
                                        (a.h.subject_offset == b.h.subject_offset &&
                                         a.h.query_offset < b.h.query_offset)));
}

// In-place heapsort of run[0, n): O(n log n) for the long runs of
// low-complexity subjects, where an insertion sort would be quadratic (a
// packed group's queries interleave in emission order).
__device__ void heap_sort_run(SegRecord* run, uint32_t n) {
    auto sift = [&](uint32_t i, uint32_t end) {
        const SegRecord r = run[i];
        for (;;) {
            uint32_t c = 2 * i + 1;
            if (c >= end) break;
            if (c + 1 < end && rec_before(run[c], run[c + 1])) ++c;
            if (!rec_before(r, run[c])) break;
            run[i] = run[c];
            i = c;
        }
        run[i] = r;
    };
    for (uint32_t i = n / 2; i-- > 0;) sift(i, n);
    for (uint32_t end = n - 1; end > 0; --end) {
        const SegRecord top = run[0];
        run[0] = run[end];
        run[end] = top;
        sift(0, end);
    }
}

constexpr uint32_t kInsertionMax = 64;  // longer runs are heap-sorted

// One (group, subject) run of records: order it by (query, subject offset,
// query offset) -- the sequential emission order within each query -- and
// move each record to its CSR position.  Runs are nearly in order already
// for one query, so short ones take an insertion sort.
__global__ __launch_bounds__(kCompactThreads) void k_gather_runs(CompactParams p) {
    if (*p.overflow) return;  // segments incomplete; the host rescans
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)p.ngroups * p.count) return;
    const uint32_t len = p.run_len[i];
    if (len == 0) return;
    const uint32_t g = (uint32_t)(i / p.count);
    SegRecord* run = p.seg + p.run_loc[i];
    if (len > kInsertionMax) {
        heap_sort_run(run, len);
    } else {
        for (uint32_t a = 1; a < len; ++a) {
            const SegRecord r = run[a];
            uint32_t m = a;
            while (m > 0 && !rec_before(run[m - 1], r)) {
                run[m] = run[m - 1];
                --m;
            }
            run[m] = r;
        }
    }
    // a run holds the records of one or more consecutive subjects (a
    // wavefront's pack), now in (subject, query, s_off, q_off) order
    const GroupInfo gi = p.groups[g];
    uint32_t cur_j = 0xFFFFFFFFu, cur_slot = 0xFFFFFFFFu, dst = 0, qid = 0;
    for (uint32_t a = 0; a < len; ++a) {
        const SegRecord r = run[a];
        if (r.j != cur_j || r.slot != cur_slot) {
            cur_j = r.j;
            cur_slot = r.slot;
            qid = p.qinfo[gi.qinfo_first + r.j].qid;
            dst = p.cnt[(uint64_t)qid * p.count + r.slot];
        }
        if (dst < p.rec_cap) {
            p.records[dst] = r.h;
            p.subjects[dst] = p.first + r.slot;
            if (p.queries) p.queries[dst] = qid;
        }
        ++dst;
    }
}

// zero-fills up to kZeroRanges arrays in one launch
__global__ void k_zero_u32(ZeroRanges z) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (int r = 0; r < z.count; ++r)
        for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < z.n[r]; i += stride) z.dst[r][i] = 0u;
}

}  // namespace

LdsPlan plan_lds(uint32_t max_table_bytes, uint32_t max_qres, uint32_t max_nq, uint32_t diag_cap, uint32_t qcap,
                 uint32_t sbuf, uint32_t claim_slots, int waves, bool table_in_lds, bool narrow) {
    auto up16 = [](uint32_t v) { return (v + 15u) & ~15u; };
    LdsPlan pl{};
    pl.max_nq = max_nq < 1 ? 1 : max_nq;
    pl.qcap = qcap;
    pl.claim_slots = claim_slots;
    pl.qres_cap = up16(max_qres < 16 ? 16 : max_qres);
    pl.wave_stride = up16(sbuf) + up16(diag_cap * (narrow ? 2u : 4u)) + claim_slots * 4 + qcap * 4 + kTaskCap * 6 +
                     kDefCap * 4 + kPack * (uint32_t)sizeof(PackSub) +
                     (pl.max_nq > 1 ? pl.max_nq * kPack * (uint32_t)sizeof(QCount) : 0u);
    uint32_t at = kCtlBytes;
    pl.table = at;
    pl.table_cap_words = table_in_lds ? up16(max_table_bytes) / 4 : 0u;
    at += pl.table_cap_words * 4;
    pl.qres = at;
    at += pl.qres_cap;
    pl.qmap = at;
    if (pl.max_nq > 1) at += pl.qres_cap;
    pl.qinfo = at;
    at += pl.max_nq * (uint32_t)sizeof(QueryInfo);
    pl.matrix = at;
    at += up16(kMatrixRows * kMatrixStride);
    pl.wave = at;
    at += pl.wave_stride * (uint32_t)waves;
    pl.total = at;
    return pl;
}

namespace {
using ScanFn = void (*)(ScanParams);
ScanFn pick_scan(int word, bool table_in_lds, bool multi, bool narrow) {
#define BPX_PICK(W, N)                                                                                \
    return table_in_lds ? (multi ? k_scan<W, true, true, N> : k_scan<W, true, false, N>)              \
                        : (multi ? k_scan<W, false, true, N> : k_scan<W, false, false, N>)
    if (word == 3) {
        if (narrow) BPX_PICK(3, true);
        BPX_PICK(3, false);
    }
    if (word == 2) {
        if (narrow) BPX_PICK(2, true);
        BPX_PICK(2, false);
    }
#undef BPX_PICK
    return nullptr;
}
}  // namespace

hipError_t prepare_scan(const ScanParams& p, bool table_in_lds, bool multi, bool narrow) {
    const ScanFn fn = pick_scan(p.word, table_in_lds, multi, narrow);
    if (!fn) return hipErrorInvalidValue;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)p.lds_total);
}

hipError_t launch_scan(const ScanParams& p, bool table_in_lds, bool multi, bool narrow, int waves, int groups,
                       hipStream_t stream, bool set_attr) {
    if (waves < 1 || waves > kMaxWaves) return hipErrorInvalidValue;
    const ScanFn fn = pick_scan(p.word, table_in_lds, multi, narrow);
    if (!fn) return hipErrorInvalidValue;
    if (set_attr) {
        const hipError_t e = prepare_scan(p, table_in_lds, multi, narrow);
        if (e != hipSuccess) return e;
    }
    const dim3 grid((unsigned)groups), block((unsigned)(64 * waves));
    hipLaunchKernelGGL(fn, grid, block, p.lds_total, stream, p);
    return hipGetLastError();
}

hipError_t launch_compact(const CompactParams& p, hipStream_t stream) {
    const uint64_t ntiles = (p.n + kTile - 1) / kTile;
    if (ntiles > 0xFFFFFFFFull) return hipErrorInvalidValue;
    if (ntiles) hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)ntiles), dim3(kCompactThreads), 0, stream, p);
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, stream, p, (uint32_t)ntiles);
    if (ntiles) hipLaunchKernelGGL(k_offsets, dim3((unsigned)ntiles), dim3(kCompactThreads), 0, stream, p);
    const uint64_t runs = (uint64_t)p.ngroups * p.count;
    if (runs) {
        const uint64_t blocks = (runs + kCompactThreads - 1) / kCompactThreads;
        hipLaunchKernelGGL(k_gather_runs, dim3((unsigned)blocks), dim3(kCompactThreads), 0, stream, p);
    }
    return hipGetLastError();
}


hipError_t launch_zero_u32(const ZeroRanges& z, hipStream_t stream) {
    uint64_t most = 0;
    for (int r = 0; r < z.count; ++r) most = most > z.n[r] ? most : z.n[r];
    if (most == 0) return hipSuccess;
    const uint64_t blocks = (most + 255) / 256;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, stream, z);
    return hipGetLastError();
}

}  // namespace bpx
