/*
 * oracle.h -- CPU restatement of blastp two-hit word finding + ungapped
 * X-drop extension.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke
 * check and the CPU-baseline leg of the benchmark link this library as the
 * checker.  The product path (libbpx.so) never calls it.
 *
 * What each function restates (reference file:line):
 *   or_lookup_build     -- the query-word neighbourhood the kernel consumes at
 *                          gpu_blastp_kernel.cu:182-192 (NCBI BlastAaLookupTable,
 *                          not vendored); restated from its definition: every
 *                          word w with sum_i M[q[p+i]][w_i] >= T, plus the exact
 *                          query word.
 *   or_scan_subject     -- word scan (gpu_blastp_kernel.cu:143-192), two-hit
 *                          diagonal state machine (:194-301, :308) and the
 *                          record/counter output (:256-276, :310-312).
 *   or_extend_two_hit   -- s_BlastAaExtendTwoHit, gpu_blastp_kernel.cu:341-415.
 *   or_extend_left      -- s_BlastAaExtendLeft,   gpu_blastp_kernel.cu:435-481.
 *   or_extend_right     -- s_BlastAaExtendRight,  gpu_blastp_kernel.cu:503-549.
 *   or_run_db           -- the per-subject loop of the CPU engine
 *                          (blast_engine.c:515-559 with the word finder of
 *                          :534-538), threaded in OID chunks of 4096 like the
 *                          NCBI chunk iterator (functions.cpp:971-972).
 *
 * Deliberately NOT restated (reference quirks): the 640-byte matrix copy
 * (gpu_cpu_common.h:28), the 2-record ring (gpu_blastp_kernel.cu:274) and the
 * uninitialised ReturnRightValue (gpu_blastp_kernel.cu:388,401).
 */
#ifndef BPX_ORACLE_H
#define BPX_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_ALPHABET 28  /* ncbistdaa, gpu_blastp_kernel.h:55 */
#define OR_CHARSIZE 5   /* bits per residue in a word index */

/* One ungapped HSP, field-for-field GPUBlastInitHitList (gpu_cpu_common.h:29-38). */
typedef struct or_hsp {
    int32_t hsp_q, hsp_s, q_off, s_off, hsp_len, score;
} or_hsp;

typedef struct or_params {
    int32_t word_length;  /* 2 or 3 */
    int32_t threshold;    /* neighbourhood threshold T */
    int32_t window;       /* two-hit window A */
    int32_t x_dropoff;    /* raw-score X-drop */
    int32_t cutoff;       /* raw-score report cutoff */
    int32_t diag_mask;    /* 2^k - 1 */
} or_params;

/* Lookup table as CSR over the 2^(5*word) word indices.  Within a cell the
 * query offsets are ascending. */
typedef struct or_lookup {
    int32_t word_length;
    uint32_t nwords;        /* 1 << (5*word_length) */
    uint32_t *cell_start;   /* nwords + 1 */
    uint16_t *entries;      /* cell_start[nwords] */
    uint32_t max_cell;      /* largest cell */
} or_lookup;

/* NCBI DiagStruct: 31-bit last_hit plus a 1-bit flag. */
typedef struct or_diag {
    int32_t last_hit;
    int32_t flag;
} or_diag;

/* Builds the table by brute force over all 28^word words at every query
 * position.  Returns 0 or -1 (bad arguments / allocation failure). */
int or_lookup_build(const uint8_t *query, int32_t qlen, const int8_t *matrix28,
                    int32_t threshold, int32_t word_length, or_lookup *out);
void or_lookup_free(or_lookup *lk);

/* Diagonal-table length for a query: next power of two >= qlen + window
 * (NCBI s_BlastDiagTableNew, not vendored; the reference passes the result
 * through at gpu_blastp.cu:63-66). */
int32_t or_diag_length(int32_t qlen, int32_t window);

typedef struct or_ext_left { int32_t maxscore, length; } or_ext_left;
typedef struct or_ext_right { int32_t maxscore, length, s_last_off; } or_ext_right;
typedef struct or_ext_two {
    int32_t hsp_q, hsp_s, hsp_len, score, right_extend, s_last_off;
} or_ext_two;

or_ext_left or_extend_left(const int8_t *matrix28, const uint8_t *subject,
                           const uint8_t *query, int32_t s_off, int32_t q_off,
                           int32_t dropoff, int32_t maxscore);
or_ext_right or_extend_right(const int8_t *matrix28, const uint8_t *subject,
                             const uint8_t *query, int32_t s_off, int32_t q_off,
                             int32_t dropoff, int32_t maxscore,
                             int32_t subject_length, int32_t query_length);
or_ext_two or_extend_two_hit(const int8_t *matrix28, const uint8_t *subject,
                             int32_t subject_length, const uint8_t *query,
                             int32_t query_length, int32_t s_left_off,
                             int32_t s_right_off, int32_t q_right_off,
                             int32_t dropoff, int32_t word_size);

/* Scans one subject.  `diag` has diag_mask+1 entries and `*diag_offset` is
 * the running NCBI diagonal offset; both persist across subjects exactly as
 * the CPU engine keeps them (offset += len + window after each subject, table
 * cleared when the offset passes INT32_MAX/4).  Records go to out[0..cap);
 * the return value is the number of records produced (may exceed cap, in
 * which case only the first cap are stored).  counters = {hits,
 * successful_extensions, hits_extended}, the d_Hits triple of
 * gpu_blastp_kernel.cu:310-312. */
int32_t or_scan_subject(const or_lookup *lk, const uint8_t *query, int32_t qlen,
                        const int8_t *matrix28, const or_params *p,
                        const uint8_t *subject, int32_t slen,
                        or_diag *diag, int32_t *diag_offset,
                        or_hsp *out, int32_t cap, int32_t counters[3]);

/* Whole-database run over subjects [0, n).  Subject i is residues[off[i] ..
 * off[i]+len[i]).  counters gets 3*n ints.  Records are returned in CSR form:
 * rec_offsets (n+1 entries) and *records (malloc'd; caller frees with
 * or_free).  nthreads <= 0 means 1.  Returns 0 or -1. */
int or_run_db(const or_lookup *lk, const uint8_t *query, int32_t qlen,
              const int8_t *matrix28, const or_params *p,
              const uint8_t *residues, const uint64_t *off, const uint32_t *len,
              uint32_t n, int nthreads, int32_t *counters,
              uint64_t *rec_offsets, or_hsp **records);
void or_free(void *ptr);

/* The 28x28 BLOSUM62 table in ncbistdaa order: data restated from the
 * reference fixture gpu_blastp.c:46-75 (all 784 bytes, not the 640 the
 * reference copies). */
const int8_t *or_blosum62(void);

#ifdef __cplusplus
}
#endif
#endif
