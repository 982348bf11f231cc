/*
 * bpx.h -- C ABI of the MI355X (gfx950) blastp word finder: two-hit word
 * finding plus ungapped X-drop extension, the work NCBI's BlastAaWordFinder
 * does for the preliminary search, run for a whole database volume per call.
 *
 * Conventions: every call returns 0 (BPX_OK) or a negative bpx_status;
 * bpx_last_error() gives the message of the last failure on this thread.
 * Handles own all device memory.  A handle is used by one host thread at a
 * time.  No entry point takes or returns framework (torch) types.
 *
 * Reference interfaces replaced (avtomaton/blastp-cuda):
 *   bpx_available        GPU_BLAST_check_availability   gpu_cpu_common.h:91, gpu_blastp.cu:336-344
 *   bpx_context_*        device selection / properties  gpu_blastp.cu:185-191, 327-329
 *   bpx_db_load          .gpu volume read + H2D          functions.cpp:620-638, gpu_blastp.cu:139,158
 *   bpx_query_prepare    lookup/PV/matrix/query staging gpu_blastp.cu:45-88, 140-145, 159-163, 177-183
 *   bpx_check_memory     GPU_BLASTP_check_memory        gpu_cpu_common.h:81-89, gpu_blastp.cu:282-334
 *   bpx_scan             GPU_BLASTP (launch)            gpu_cpu_common.h:44-63, gpu_blastp.cu:16-222
 *   bpx_batch_*, bpx_scan_batch: several queries per launch -- the reference's
 *                        multi-context launch (num_queries contexts,
 *                        gpu_blastp_kernel.cu:238-241), with per-query
 *                        diagonal state
 *   bpx_result_wait      GPU_BLASTP_get_data (sync)     gpu_cpu_common.h:65-69, gpu_blastp.cu:224-259
 *   bpx_result_*         h_Hits / h_GPUBlastInitHitList consumed at blast_engine.c:520-559
 *   bpx_*_destroy        GPU_BLASTP_free_memory         gpu_cpu_common.h:71-79, gpu_blastp.cu:262-280
 */
#ifndef BPX_H
#define BPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BPX_ALPHABET 28 /* ncbistdaa codes 0..27 */

typedef enum bpx_status {
    BPX_OK = 0,
    BPX_E_ARG = -1,       /* bad argument */
    BPX_E_NODEV = -2,     /* no usable GPU */
    BPX_E_HIP = -3,       /* HIP runtime error */
    BPX_E_NOMEM = -4,     /* device or host allocation failed */
    BPX_E_BUDGET = -5,    /* exceeds the device memory / LDS budget */
    BPX_E_STATE = -6      /* call out of order (e.g. result read before wait) */
} bpx_status;

typedef struct bpx_ctx bpx_ctx;
typedef struct bpx_db bpx_db;
typedef struct bpx_query bpx_query;
typedef struct bpx_batch bpx_batch;
typedef struct bpx_result bpx_result;

/* Word-finding parameters of one query (one NCBI "context").  Mirrors what
 * the reference reads from the lookup table, diagonal table and
 * BlastInitialWordParameters (gpu_blastp.cu:63-82):
 *   word_length  lookup->word_length (2 or 3)
 *   threshold    neighbourhood threshold T of the lookup table
 *   window       ewp->diag_table->window (two-hit window A)
 *   x_dropoff    word_params->cutoffs[ctx].x_dropoff (raw score)
 *   cutoff_score word_params->cutoffs[ctx].cutoff_score (raw score)
 *   diag_mask    ewp->diag_table->diag_mask; 0 = derive it as
 *                (next power of two >= qlen + window) - 1 */
typedef struct bpx_word_params {
    int32_t word_length;
    int32_t threshold;
    int32_t window;
    int32_t x_dropoff;
    int32_t cutoff_score;
    int32_t diag_mask;
} bpx_word_params;

/* One ungapped HSP; field-for-field GPUBlastInitHitList (gpu_cpu_common.h:29-38),
 * i.e. the arguments of BlastSaveInitHsp (blast_engine.c:545-551). */
typedef struct bpx_init_hsp {
    int32_t hsp_q;
    int32_t hsp_s;
    int32_t query_offset;
    int32_t subject_offset;
    int32_t hsp_len;
    int32_t score;
} bpx_init_hsp;

/* Per (query, subject) counters: the d_Hits triple of the reference
 * (gpu_blastp_kernel.cu:310-312), read at blast_engine.c:520-524. */
typedef struct bpx_subject_stats {
    int32_t hits;                  /* lookup hits (word matches) */
    int32_t successful_extensions; /* records saved (score >= cutoff) */
    int32_t hits_extended;         /* two-hit extensions attempted */
} bpx_subject_stats;

/* Sums Blast_UngappedStatsUpdate (blast_engine.c:553-556) would accumulate
 * over every subject of one query. */
typedef struct bpx_ungapped_stats {
    int64_t lookup_hits;
    int64_t num_seqs_lookup_hits;
    int64_t init_extends;
    int64_t good_init_extends;
    int64_t num_seqs_passed;
} bpx_ungapped_stats;

/* Launch configuration of the scan kernels (the reference's -gpu_blocks /
 * -gpu_threads, cmdline_flags.cpp:218-222, re-expressed for gfx950).  Any
 * field left 0 takes the library default or the autotuned value. */
typedef struct bpx_launch {
    int32_t waves_per_group;    /* wavefronts per workgroup (1..16); 0 = the most that fit */
    int32_t groups_per_cu;      /* resident workgroups per CU (grid = CUs x this) */
    int32_t subjects_per_item;  /* subjects per work item a workgroup claims; 0 = subjects claimed
                                   by wavefronts device-wide, no items */
    int32_t queue_capacity;     /* per-wave LDS hit ring entries (power of two) */
    int32_t subject_buffer;     /* per-wave LDS subject buffer bytes (power of two >= 1024) */
    int32_t claims_per_share;   /* device-wide claims: a wavefront claims about 1/claims_per_share
                                   of its share of an item's subjects at a time (at most 16
                                   subjects); 0 = 16 */
} bpx_launch;

/* How a batch packs its queries into scan groups: queries shorter than
 * group_residues are concatenated (at most group_queries per group, and while
 * their diagonal-table lengths sum to at most group_diag_slots) and share one
 * lookup table and one pass over every subject; longer queries form groups
 * of one.  A group's summed diagonal-table length is the diagonal table each
 * wavefront holds in LDS while it scans the group, so group_diag_slots bounds
 * that table and with it how many wavefronts fit a CU.  Results do not
 * depend on the packing. */
typedef struct bpx_batch_policy {
    int32_t group_residues;   /* 0 = default */
    int32_t group_queries;    /* 0 = default, at most 255 */
    int32_t group_diag_slots; /* 0 = no bound beyond the two above */
} bpx_batch_policy;

/* Output selection for a scan. */
#define BPX_OUT_SUBJECT_STATS 1u /* per-subject counters to the host (the kernel writes them to pinned memory) */
#define BPX_OUT_GRAPH 2u         /* capture the scan's device work in a HIP graph (bpx_result_replay) */
#define BPX_OUT_PHASES 4u        /* per-wavefront phase times of the scan kernels (bpx_result_phases) */

/* ---- device / context ------------------------------------------------- */
int bpx_available(void);
int bpx_context_create(int device, bpx_ctx **out);
int bpx_context_destroy(bpx_ctx *ctx);
const char *bpx_last_error(void);
/* Number of compute units and usable LDS bytes per workgroup of the device. */
int bpx_context_info(const bpx_ctx *ctx, int32_t *num_cus, int64_t *lds_bytes,
                     int64_t *hbm_bytes);

/* ---- database: a resident int8 residue arena ---------------------------
 * residues: ncbistdaa codes; subject i is residues[offsets[i] ..
 * offsets[i]+lengths[i]).  The library repacks it into its own aligned
 * arena on the device (with up to 16 host threads, through pinned staging
 * buffers of its own) and checks every code; the host arrays may be freed
 * on return. */
int bpx_db_load(bpx_ctx *ctx, const uint8_t *residues, const uint64_t *offsets,
                const uint32_t *lengths, uint32_t num_subjects, bpx_db **out);
int bpx_db_info(const bpx_db *db, uint32_t *num_subjects, uint64_t *num_residues,
                uint64_t *arena_bytes);
int bpx_db_destroy(bpx_db *db);

/* ---- query: lookup table built on the host, uploaded once ---------------
 * query: ncbistdaa codes, qlen <= 65535.  matrix: 28x28 int8 row-major
 * (row = query residue) or NULL for the built-in BLOSUM62. */
int bpx_query_prepare(bpx_ctx *ctx, const uint8_t *query, int32_t qlen,
                      const int8_t *matrix, const bpx_word_params *params,
                      bpx_query **out);
int bpx_query_info(const bpx_query *q, int64_t *table_bytes, int32_t *lds_resident,
                   int32_t *num_present_words, int64_t *num_entries);
int bpx_query_destroy(bpx_query *q);

/* Device bytes a scan of subjects [first, first+count) of db would need and
 * the LDS bytes per workgroup; returns BPX_E_BUDGET if it does not fit the
 * device (or the context's limit, bpx_context_limit). */
int bpx_check_memory(bpx_ctx *ctx, const bpx_db *db, const bpx_query *q,
                     uint32_t first, uint32_t count, const bpx_launch *launch,
                     size_t *device_bytes, size_t *lds_bytes);
/* Caps the device bytes bpx_check_memory (and the server's default result
 * budget) plans with, below the device's own size -- for a GPU shared with
 * other work.  0 restores the device size.  The check of GPU_BLASTP_check_memory
 * (gpu_blastp.cu:316-333) against totalGlobalMem, with the total settable. */
int bpx_context_limit(bpx_ctx *ctx, int64_t hbm_bytes);
/* Frees the device and pinned host buffers the context keeps cached for
 * later scans (results and queries hold theirs until destroyed), after the
 * context's streams finish.  A context releases its cache by itself when an
 * allocation fails; this returns the memory at once, for a GPU shared with
 * other work.  The counterpart of GPU_BLASTP_free_memory
 * (gpu_cpu_common.h:71), which frees a search's buffers after each volume;
 * here buffers are reused across scans until trimmed. */
int bpx_context_trim(bpx_ctx *ctx);
/* Footprint of a query of qlen residues scanned alone against db, planned
 * without building its table (the table read from HBM, the placement every
 * query can take): its diagonal slots, the most wavefronts per workgroup it
 * runs at and the LDS bytes per workgroup there.  BPX_E_BUDGET if it cannot
 * run at all (more than 65,536 diagonal slots, or not even one wavefront's
 * buffers fit the LDS) -- the shared-memory and diagonal-array checks of
 * gpu_blastp.cu:63-65, 184-191, made before any work is queued.  params as
 * for bpx_query_prepare (word_length 0: blastp defaults); launch nullable. */
int bpx_query_footprint(bpx_ctx *ctx, const bpx_db *db, int32_t qlen, const bpx_word_params *params,
                        const bpx_launch *launch, uint32_t *diag_slots, int32_t *waves,
                        size_t *lds_bytes);

/* ---- batch: many queries scanned in one persistent launch ---------------
 * queries[i] has qlens[i] residues.  params gives word_length, threshold,
 * window (shared by the batch) and the default x_dropoff / cutoff_score;
 * x_dropoff / cutoff (nullable) override them per query (per NCBI context,
 * gpu_blastp.cu:77-82).  params->diag_mask must be 0 (derived per query). */
int bpx_batch_prepare(bpx_ctx *ctx, uint32_t num_queries, const uint8_t *const *queries,
                      const int32_t *qlens, const int8_t *matrix,
                      const bpx_word_params *params, const int32_t *x_dropoff,
                      const int32_t *cutoff, const bpx_batch_policy *policy,
                      bpx_batch **out);
int bpx_batch_info(const bpx_batch *b, uint32_t *num_queries, uint32_t *num_groups,
                   int64_t *table_bytes);
int bpx_batch_destroy(bpx_batch *b);

/* ---- scan ---------------------------------------------------------------
 * Asynchronous: enqueues the scan of subjects [first, first+count) on the
 * context's stream and returns.  launch may be NULL (defaults). */
int bpx_scan(bpx_ctx *ctx, const bpx_query *q, const bpx_db *db, uint32_t first,
             uint32_t count, const bpx_launch *launch, uint32_t out_flags,
             bpx_result **out);
/* The same for every query of a batch, in one persistent launch per table
 * placement.  Results are identical to scanning each query alone. */
int bpx_scan_batch(bpx_ctx *ctx, const bpx_batch *b, const bpx_db *db, uint32_t first,
                   uint32_t count, const bpx_launch *launch, uint32_t out_flags,
                   bpx_result **out);
/* Blocks until the scan and its result copies are complete. */
int bpx_result_wait(bpx_result *r);
/* 1 if the scan's device work has finished (bpx_result_wait will not block
 * on the device), 0 if not, < 0 on error.  Never blocks. */
int bpx_result_ready(const bpx_result *r);
/* Times of the last scan, measured with HIP events on the stream the
 * kernels run on: the word-finder kernels alone, and the whole device phase
 * (word finder + compaction + copies to the host). */
int bpx_result_timing(const bpx_result *r, float *kernel_ms, float *device_ms);
/* Number of records (all queries and subjects) and the ungapped stats
 * summed over the scan's queries. */
int bpx_result_summary(const bpx_result *r, uint64_t *num_records,
                       bpx_ungapped_stats *stats);
/* Per-query ungapped stats (num_queries entries, in batch order). */
int bpx_result_query_stats(const bpx_result *r, bpx_ungapped_stats *stats,
                           uint32_t capacity);
/* Records, ordered by (query, subject, subject_offset, query_offset) -- for
 * one query and subject that is the sequential emission order; subjects[i]
 * and queries[i] are the absolute subject index and the batch query index of
 * records[i].  Any output pointer may be NULL. */
int bpx_result_records(const bpx_result *r, bpx_init_hsp *records, uint32_t *subjects,
                       uint32_t *queries, uint64_t capacity);
/* Per-(query, subject) counters, query-major (num_queries x count entries);
 * available when the scan was made with BPX_OUT_SUBJECT_STATS. */
int bpx_result_subject_stats(const bpx_result *r, bpx_subject_stats *stats,
                             uint64_t capacity);
/* Graph mode (scan made with BPX_OUT_GRAPH): re-runs the captured device
 * work of the scan -- counter resets, every launch class, compaction, copies
 * -- as one hipGraphLaunch on the same buffers, for the same queries,
 * database and range.  Call after bpx_result_wait; wait again before reading.
 * Timing then covers the whole graph.  No reference counterpart: the
 * reference re-issues every copy and launch per search (gpu_blastp.cu:139-201). */
int bpx_result_replay(bpx_result *r);
/* Phase times of a scan made with BPX_OUT_PHASES, from each wavefront's
 * records on the device's 100 MHz real-time clock, summed over the scan's
 * launch classes.  No reference counterpart (its only timer is around the
 * launch, functions.cpp:683-695).  stage_ms_mean is the mean time a
 * wavefront spent staging subjects from HBM into LDS (descriptor reads to
 * residues in LDS); bytes_staged / stage_ms_mean is the rate of that
 * streaming phase while it runs.  group_ms_mean is the mean time at group
 * changes (waiting for the workgroup's other wavefronts, restaging the
 * group's table and residues).  flush_ms_mean is the mean time in
 * extension flushes (the queued two-hit X-drop extensions and the writing
 * of their records); extensions counts the extensions those flushes ran.
 * wave_ms_* are wavefront lifetimes; span_ms runs from the first start to
 * the last end of each launch class, summed.  busy_fraction is the summed
 * wavefront lifetimes over the wavefront slots of the launch classes
 * (wavefronts x span of each): 1 - busy_fraction is the share of the
 * launches' time that finished wavefronts sat idle waiting for the slowest,
 * which bounds what any reordering of the work could save. */
typedef struct bpx_phase_stats {
    uint32_t waves;
    uint32_t classes;
    double span_ms;
    double wave_ms_mean;
    double wave_ms_min;
    double stage_ms_mean;
    double group_ms_mean;
    uint64_t subjects_staged;
    uint64_t bytes_staged;
    double flush_ms_mean;
    uint64_t extensions;
    double busy_fraction;
} bpx_phase_stats;
int bpx_result_phases(const bpx_result *r, bpx_phase_stats *out);
int bpx_result_destroy(bpx_result *r);

/* ---- scan without allocation ---------------------------------------------
 * No counterpart in the reference (which allocates its device buffers on
 * every search, gpu_blastp.cu:139-175).  For callers that cannot allocate
 * memory while they scan: tight serving loops, harnesses that forbid
 * hipMalloc.  The scan of bpx_scan for one prepared query over the whole
 * database, synchronous, with every buffer beyond its inputs and outputs in
 * the library's fixed scratch array: one per GPU the library runs on
 * (BPX_NOALLOC_SCRATCH_BYTES, held until the process ends).  Limits: at most
 * BPX_NOALLOC_MAX_SUBJECTS subjects; BPX_NOALLOC_SEG_RECORDS record slots,
 * split evenly among the launch's wavefronts -- a wavefront that saves more
 * records than its share fails the call with BPX_E_BUDGET (bpx_scan sizes
 * its buffers instead).  Outputs: d_records / d_subjects are device buffers
 * of record_capacity entries (bpx_result_records' records and subject ids,
 * in the same order); d_stats (nullable) a device buffer of one
 * bpx_subject_stats per subject; *num_records and *stats (nullable) on the
 * host.  When the records exceed record_capacity, none past it are written,
 * *num_records still receives their number and the call returns BPX_E_ARG.
 * No hipMalloc, no pinned or pageable host buffer, no allocating library
 * call.  The scratch keeps the last call's contents (every call clears what
 * must start at zero); calls that overlap on one GPU -- from other threads
 * or other contexts of that device -- share it and race. */
#define BPX_NOALLOC_MAX_SUBJECTS 2097152u
#define BPX_NOALLOC_SEG_RECORDS 1048576u
#define BPX_NOALLOC_SCRATCH_BYTES 58729984u
int bpx_scan_noalloc(bpx_ctx *ctx, const bpx_query *q, const bpx_db *db, bpx_init_hsp *d_records,
                     uint32_t *d_subjects, uint64_t record_capacity, bpx_subject_stats *d_stats,
                     uint64_t *num_records, bpx_ungapped_stats *stats);

/* ---- autotuner ----------------------------------------------------------
 * Times batched scans of subjects [sample_first, sample_first+sample_count)
 * under candidate packing policies and launch shapes and returns the
 * fastest.  No reference counterpart: the reference takes its launch shape
 * from -gpu_blocks / -gpu_threads (cmdline_flags.cpp:218-222).  Coordinate
 * search, each step keeping the best so far: packing policy -- residues per
 * group, queries per group, then a bound on each group's summed diagonal
 * slots (the per-wavefront LDS diagonal table; these set how many wavefronts
 * fit a CU and the table placement) -- then subject buffer, hit-ring
 * capacity, residues once more, then subjects per wavefront claim
 * (claims_per_share), the last
 * timed on twice the sample (or the whole database) since the claim
 * size follows the number of subjects.  The arguments before db are those of
 * bpx_batch_prepare.  trials (nullable) receives every configuration timed.
 * With the environment variable BPX_TUNE_LOG set (and not "0") each timed
 * candidate also prints one line on stderr as it finishes. */
typedef struct bpx_tune_trial {
    bpx_batch_policy policy;
    bpx_launch launch;
    float sample_ms;            /* word-finder kernel time on the sample */
} bpx_tune_trial;
typedef struct bpx_tune_result {
    bpx_batch_policy policy;
    bpx_launch launch;
    float sample_ms;
    int32_t candidates;
} bpx_tune_result;
int bpx_autotune_batch(bpx_ctx *ctx, uint32_t num_queries, const uint8_t *const *queries,
                       const int32_t *qlens, const int8_t *matrix, const bpx_word_params *params,
                       const int32_t *x_dropoff, const int32_t *cutoff, const bpx_db *db,
                       uint32_t sample_first, uint32_t sample_count, bpx_tune_result *out,
                       bpx_tune_trial *trials, uint32_t trial_capacity, uint32_t *num_trials);

/* ---- learned cost model --------------------------------------------------
 * No reference counterpart (the reference orders subjects by length only,
 * functions.cpp:125-149).  A 32-64-64-1 ReLU MLP with bf16 weights and
 * activations and fp32 accumulation, trained and run on the GPU by the
 * library's own MFMA kernels.  It predicts per-subject scan work from
 * per-subject features and is used only to order and bin work
 * (bpx_db_set_cost); results never depend on it. */
#define BPX_CM_FEATURES 32  /* feature vector width (22 used, rest zero) */
#define BPX_CM_PARAMS 7312  /* W1[32][64] b1[64] W2[64][64] b2[64] W3[64][16] b3[16] */
typedef struct bpx_costmodel bpx_costmodel;
/* Features of every subject, computed on the GPU, [num_subjects x 32] fp32
 * row-major: log(1 + length), the composition fractions of the 20 standard
 * residues (ARNDCQEGHILKMFPSTWYV), log(1 + sum over the subject's words of
 * word_counts[word]) (a sampled query-word hit count), zeros. */
int bpx_db_features(bpx_ctx *ctx, const bpx_db *db, const uint16_t *word_counts,
                    int32_t word_length, float *features);
int bpx_costmodel_create(int device, uint64_t seed, bpx_costmodel **out);
/* get (set = 0) or set the fp32 parameters (BPX_CM_PARAMS floats) */
int bpx_costmodel_params(bpx_costmodel *cm, float *params, int set);
/* Adam on mean squared error; epoch_loss (nullable) gets each epoch's mean loss. */
int bpx_costmodel_train(bpx_costmodel *cm, const float *features, const float *target,
                        uint32_t n, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                        float *epoch_loss);
int bpx_costmodel_predict(bpx_costmodel *cm, const float *features, uint32_t n, float *out);
/* Mean squared error and its gradient (BPX_CM_PARAMS floats) over n rows. */
int bpx_costmodel_grad(bpx_costmodel *cm, const float *features, const float *target,
                       uint32_t n, float *loss, float *grad);
const char *bpx_costmodel_last_error(void);
int bpx_costmodel_destroy(bpx_costmodel *cm);
/* Processing order by predicted cost: subjects of length >= 15,360 first
 * (longest first), then the rest by descending cost.  The order is cut into
 * 8 bins of equal predicted work; a scan of the whole database gives each
 * bin its own work item, whose wavefronts claim subjects in chunks sized by
 * the bin's mean cost.  cost = NULL restores the default longest-first
 * order without bins.  Scans with a window above 1,024 are refused while a
 * cost order is set.  Not while a scan of this database is in flight. */
int bpx_db_set_cost(bpx_db *db, const float *cost);

/* ---- query server -------------------------------------------------------
 * No reference counterpart (blastp-cuda serves one search per process).
 * Continuous batching over a resident database: submitted queries wait in a
 * queue; a worker thread admits them into batched scans when a device slot
 * is free (at most max_inflight batches, on two stream lanes), the batch's
 * result buffers fit hbm_budget and its queries' diagonal tables fit
 * diag_budget -- at once when the device is idle, otherwise when max_batch
 * queries wait or the oldest has waited max_wait_us.  A query that cannot
 * run against the database at all (bpx_query_footprint) is refused by
 * bpx_server_submit with BPX_E_BUDGET instead of failing its batch.  Each request's records and stats are exactly those of
 * scanning it alone.  While the server runs, its worker thread owns the
 * context: make no other calls on that context from other threads. */
typedef struct bpx_server bpx_server;
typedef struct bpx_server_config {
    int32_t max_batch;          /* queries per batch (0 = 256) */
    int32_t max_inflight;       /* batches on the device at once, 1..2 (0 = 2) */
    int32_t max_wait_us;        /* gathering time while a batch runs (0 = 10,000) */
    int32_t diag_budget;        /* diagonal-table slots (summed over its queries) one batch may
                                   take (0 = 1,048,576); a query above it runs in a batch alone */
    int64_t hbm_budget;         /* result-buffer bytes one batch may take (0 = 1/8 of HBM) */
    bpx_word_params params;     /* word_length 0 = blastp defaults; diag_mask must be 0 */
    bpx_batch_policy policy;
    bpx_launch launch;
} bpx_server_config;
typedef struct bpx_request_info {
    uint64_t num_records;
    bpx_ungapped_stats stats;
    double latency_ms;          /* submit -> result ready */
    double queue_ms;            /* submit -> its batch launched */
    uint32_t batch_size;        /* queries in its batch */
    uint32_t reserved;
} bpx_request_info;
int bpx_server_create(bpx_ctx *ctx, const bpx_db *db, const bpx_server_config *cfg, bpx_server **out);
int bpx_server_submit(bpx_server *s, const uint8_t *query, int32_t qlen, uint64_t *ticket);
/* Waits up to timeout_ms (-1 = no limit; BPX_E_STATE on timeout); returns
 * the request's own status. */
int bpx_server_wait(bpx_server *s, uint64_t ticket, int32_t timeout_ms, bpx_request_info *info);
/* Copies the request's records (ordered by subject, subject offset, query
 * offset) and subject ids, then forgets the ticket. */
int bpx_server_records(bpx_server *s, uint64_t ticket, bpx_init_hsp *records, uint32_t *subjects,
                       uint64_t capacity);
int bpx_server_counters(bpx_server *s, uint64_t *batches, uint64_t *completed);
/* Finishes every queued request, then stops the worker. */
int bpx_server_destroy(bpx_server *s);

/* ---- NCBI protein database volumes (format version 4) --------------------
 * base is a path without extension.  bpx_volume_read reads base.pin and
 * base.psq: the index header walk follows the reference's reader
 * (functions.cpp:181-215: version, type == 1 (protein), title, date, number
 * of sequences), then the residue count, longest length and the sequence
 * offsets; sequence i is [off[i], off[i+1] - 1) of .psq.  bpx_db_volumes
 * lists a database's volumes -- the DBLIST line of base.pal
 * (functions.cpp:217-353) or base itself.  bpx_db_load_volumes loads them
 * all, in order, as one resident arena (subject i = OID i).
 * bpx_volume_write writes the same layout (empty .phr headers).  Host-only
 * except bpx_db_load_volumes. */
typedef struct bpx_volume bpx_volume;
int bpx_volume_read(const char *base, bpx_volume **out);
int bpx_volume_info(const bpx_volume *v, uint32_t *num_sequences, uint64_t *num_residues,
                    uint32_t *max_length, const char **title);
int bpx_volume_arrays(const bpx_volume *v, const uint8_t **residues, const uint64_t **offsets,
                      const uint32_t **lengths);
int bpx_volume_destroy(bpx_volume *v);
int bpx_volume_write(const char *base, const char *title, const uint8_t *residues,
                     const uint64_t *offsets, const uint32_t *lengths, uint32_t n);
int bpx_db_volumes(const char *base, char *names, size_t capacity, size_t *needed,
                   uint32_t *count);
int bpx_db_load_volumes(bpx_ctx *ctx, const char *base, bpx_db **out);
const char *bpx_volume_last_error(void);

/* ---- the reference's own GPU database files (.gpu / .gpuinfo) ------------
 * Replaces ReadGPU_BLASTP_database + Read_GPU_information_file(_preliminary)
 * (functions.cpp:47-179, 618-639), which read the files that
 * CreateGPU_BLASTP_database (functions.cpp:364-613) writes: sequences in
 * length order, interleaved 4 residues at a time within groups of
 * blocks * threads sequences.  bpx_gpu_info reads the .gpuinfo header;
 * bpx_gpu_volume_read de-interleaves volume `volume` into a bpx_volume
 * (sequence i = OID break-limit-of-the-previous-volume + i; lengths from the
 * last non-zero residue, since the files hold none).  The .gpu file is
 * opened by the name the .gpuinfo records, else by that file name next to
 * the .gpuinfo.  Host-only. */
int bpx_gpu_info(const char *gpuinfo, uint32_t *num_volumes, uint32_t *group_size, uint32_t *percentage);
int bpx_gpu_volume_read(const char *gpuinfo, uint32_t volume, bpx_volume **out);

/* The built-in BLOSUM62 table (784 bytes, ncbistdaa order). */
const int8_t *bpx_blosum62(void);

/* Host-only (no device needed): the lookup table bpx_query_prepare builds,
 * unpacked to CSR over all 2^(5*word_length) word indices -- the query
 * offsets listed under word w are entries[cell_start[w] .. cell_start[w+1]),
 * ascending.  Replaces NCBI BlastAaLookupTableNew + BlastAaLookupIndexQuery
 * (not vendored), whose result the reference reads at gpu_blastp.cu:45-61.
 * cell_start needs 2^(5*word_length)+1 entries; *num_entries receives the
 * total even when capacity is too small (then BPX_E_ARG). */
int bpx_table_build(const uint8_t *query, int32_t qlen, const int8_t *matrix,
                    int32_t threshold, int32_t word_length, uint32_t *cell_start,
                    uint16_t *entries, uint64_t capacity, uint64_t *num_entries);

#ifdef __cplusplus
}
#endif
#endif /* BPX_H */
