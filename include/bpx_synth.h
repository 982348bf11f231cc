/*
 * bpx_synth.h -- deterministic synthetic protein workloads for the word
 * finder: queries and databases with planted homologs and compositionally
 * biased subjects.  No counterpart in the reference (which runs on a real
 * database, README:383-384); this generator exists so the GPU path and the
 * CPU oracle can be driven with identical, seeded inputs of a stated shape.
 *
 * Every random choice is a function of (seed, index) only -- the same seed
 * gives the same bytes on every machine and with any thread count -- and
 * bpx_synth_db_sha256 fingerprints the result.  Residues are ncbistdaa codes
 * of the 20 standard amino acids, drawn from BLOSUM62 background frequencies.
 */
#ifndef BPX_SYNTH_H
#define BPX_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bpx_synth_db_spec {
    uint64_t seed;
    uint32_t num_subjects;
    /* background subject lengths: lognormal(median, sigma) clipped */
    double len_median;
    double len_sigma;
    uint32_t len_min;
    uint32_t len_max;
    /* planted homologs of the queries: exact count if homolog_count > 0,
     * else a fraction of the subjects */
    uint32_t homolog_count;
    double homolog_fraction;
    double identity_min, identity_max; /* per-homolog identity, uniform */
    double indel_min, indel_max;       /* per-position indel rate, uniform */
    /* compositionally biased subjects (Dirichlet composition + a tandem
     * repeat of a motif), as a fraction of the non-homolog subjects */
    double biased_fraction;
    double dirichlet_alpha;
    uint32_t motif_min, motif_max;
} bpx_synth_db_spec;

typedef struct bpx_synth_db bpx_synth_db;

/* Background query of `len` residues for index `index` under `seed`. */
int bpx_synth_query(uint64_t seed, uint32_t index, uint32_t len, uint8_t *out);
/* n query lengths: log-uniform in [lo, hi] (kind 0) or lognormal with the
 * given median/sigma clipped to [lo, hi] (kind 1). */
int bpx_synth_query_lengths(uint64_t seed, uint32_t n, int kind, double lo,
                            double hi, double median, double sigma,
                            uint32_t *out);
/* Generates a database.  queries[i] (qlens[i] residues) are the homolog
 * sources; nthreads <= 0 uses all hardware threads. */
int bpx_synth_db_generate(const bpx_synth_db_spec *spec,
                          const uint8_t *const *queries, const uint32_t *qlens,
                          uint32_t num_queries, int nthreads,
                          bpx_synth_db **out);
/* Borrowed views of the generated arrays (valid until free).  Subjects are
 * packed back to back: offsets[i+1] = offsets[i] + lengths[i]. */
int bpx_synth_db_arrays(const bpx_synth_db *db, const uint8_t **residues,
                        const uint64_t **offsets, const uint32_t **lengths,
                        uint32_t *num_subjects, uint64_t *num_residues);
/* Per-subject kind: 0 background, 1 planted homolog, 2 biased. */
int bpx_synth_db_kinds(const bpx_synth_db *db, const uint8_t **kinds);
/* SHA-256 over lengths (u32 little-endian) then residues; 64 hex chars + NUL. */
int bpx_synth_db_sha256(const bpx_synth_db *db, char hex[65]);
int bpx_synth_db_free(bpx_synth_db *db);
/* SHA-256 of an arbitrary buffer (64 hex chars + NUL). */
int bpx_sha256_hex(const void *data, uint64_t len, char hex[65]);

#ifdef __cplusplus
}
#endif
#endif /* BPX_SYNTH_H */
