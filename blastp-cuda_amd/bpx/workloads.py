"""The five seeded synthetic workloads the word finder is measured and
checked on.  bench.py and the tests build their inputs here, so both see the
same bytes (each database is fingerprinted by bpx_synth_db_sha256).

1. plain, CPU-sized: one 300-aa query against 10,000 subjects (lognormal
   lengths, median 300, sigma 0.6, clipped 30-5,000) with 50 planted homologs
   at 50 % identity and 2 % indels.
2. plain at scale: one 500-aa query against 2,000,000 subjects (same
   lengths), 0.5 % planted homologs at 30-70 % identity, 0-3 % indels.
3. batched: 1,024 queries of log-uniform length 50-2,000 against 500,000
   subjects, 1 % planted homologs of random queries at 40-60 % identity.
4. skewed: the same 1,024 queries against 500,000 subjects of lognormal
   length (median 350, sigma 1.2, clipped 20-35,000); 20 % compositionally
   biased (Dirichlet alpha 0.3 composition plus tandem repeats of 5-30-aa
   motifs).
5. serving: single queries of lognormal length (median 350, sigma 0.6,
   clipped 20-3,000) arriving as a Poisson stream against the database of
   workload 2.

Scoring everywhere: BLOSUM62, word 3, T = 11, window 40, x_dropoff 16,
cutoff 41 (bpx.default_params).
"""
from __future__ import annotations

from . import synth_db, synth_query, synth_query_lengths

CONFIG1_SPEC = dict(seed=0, num_subjects=10_000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                    homolog_count=50, identity_min=0.5, identity_max=0.5, indel_min=0.02, indel_max=0.02)
CONFIG2_SPEC = dict(seed=2, num_subjects=2_000_000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                    homolog_fraction=0.005, identity_min=0.3, identity_max=0.7, indel_min=0.0, indel_max=0.03)
CONFIG3_SPEC = dict(seed=3, num_subjects=500_000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                    homolog_fraction=0.01, identity_min=0.4, identity_max=0.6, indel_min=0.0, indel_max=0.03)
CONFIG4_SPEC = dict(seed=4, num_subjects=500_000, len_median=350, len_sigma=1.2, len_min=20, len_max=35000,
                    homolog_fraction=0.01, identity_min=0.4, identity_max=0.6, indel_min=0.0, indel_max=0.03,
                    biased_fraction=0.2, dirichlet_alpha=0.3, motif_min=5, motif_max=30)
# serving query lengths: the configuration gives no sigma; 0.6 (that of the
# database lengths) is used
SERVING_SIGMA = 0.6

# SHA-256 of each database (lengths as u32 little-endian, then residues)
SHA256 = {
    1: "9ca80c0274d4e6a0bc1b189d0a290a3a1509b02c38792ae13e3fd1c673b130b9",
    2: "8d43a5d40e9e6287ab2e9f4a9f5602cede04eb5f0e4ed81eef5100cff4130a62",
    3: "0870d2b97a2296dfe18f03734b725b70a3a7d7b5c882dfebb6b2e73cf6c8d382",
    4: "5dca718e7308923b188b463573afeca29de38b36b6d530ff692326d00dafe0e5",
}


def config1(threads: int = 0):
    q = synth_query(0, 0, 300)
    return q, synth_db(CONFIG1_SPEC, [q], nthreads=threads)


def config2(threads: int = 0):
    q = synth_query(0, 0, 500)
    return q, synth_db(CONFIG2_SPEC, [q], nthreads=threads)


def batch_queries():
    """The 1,024 queries of workloads 3 and 4 (log-uniform 50-2,000 aa)."""
    lens = synth_query_lengths(3, 1024, "loguniform", 50, 2000)
    return [synth_query(3, i, int(n)) for i, n in enumerate(lens)]


def config3(threads: int = 0):
    qs = batch_queries()
    return qs, synth_db(CONFIG3_SPEC, qs, nthreads=threads)


def config4(threads: int = 0):
    qs = batch_queries()
    return qs, synth_db(CONFIG4_SPEC, qs, nthreads=threads)


def serving_queries(n: int, seed: int = 5):
    """n serving queries (lognormal lengths, median 350, clipped 20-3,000)."""
    lens = synth_query_lengths(seed, n, "lognormal", 20, 3000, 350, SERVING_SIGMA)
    return [synth_query(seed, i, int(L)) for i, L in enumerate(lens)]
