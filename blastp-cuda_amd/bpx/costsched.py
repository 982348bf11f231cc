"""Cost-model scheduling of batched scans.

A scan's work per subject is driven by its lookup hits and two-hit
extensions, which depend on composition and repeats as much as on length.
The cost model (bpx.CostModel, GPU-trained MFMA MLP) learns
log(1 + hits + KAPPA * extensions) per subject -- summed over a fixed sample
of the batch's queries -- from the per-subject counters the plain kernel logs
(bpx_result_subject_stats), over seeded training databases.  Its prediction,
plus a per-residue term, orders and bins the database's work
(bpx.db_set_cost).  It never filters: results are the same in any order.
"""
import numpy as np

from . import (Context, CostModel, db_features, db_set_cost, scan, synth_db, word_counts)

KAPPA = 24.0   # work of one two-hit extension in lookup-hit units (~27 X-drop steps)
ALPHA = 0.6    # work of one residue of one scan pass in lookup-hit units


def sample_queries(queries, n=16):
    """Every len/n-th query: the fixed sample the features and labels use."""
    step = max(1, len(queries) // n)
    return [q for q in queries[::step]][:n]


def work_labels(ctx: Context, db, queries):
    """log1p(hits + KAPPA * extensions) per subject, summed over `queries`,
    from plain scans with per-subject counters."""
    hits = np.zeros(db.num_subjects, dtype=np.float64)
    ext = np.zeros(db.num_subjects, dtype=np.float64)
    for q in queries:
        dq = ctx.prepare_query(q)
        r = scan(ctx, dq, db, subject_stats=True).wait()
        st = r.subject_stats()
        hits += st["hits"]
        ext += st["hits_extended"]
        r.close()
        dq.close()
    return np.log1p(hits + KAPPA * ext).astype(np.float32)


def training_set(ctx: Context, spec: dict, queries, counts, batches=100, subjects=2000, seed0=1000):
    """Features and labels over `batches` seeded databases of `subjects`
    subjects drawn like `spec` (homologs planted from `queries`)."""
    Xs, ys = [], []
    for b in range(batches):
        s = dict(spec, seed=seed0 + b, num_subjects=subjects)
        s.pop("homolog_count", None)
        sdb = synth_db(s, queries, nthreads=16)
        ddb = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
        Xs.append(db_features(ctx, ddb, counts))
        ys.append(work_labels(ctx, ddb, queries))
        ddb.close()
    return np.concatenate(Xs), np.concatenate(ys)


def fit(ctx: Context, spec: dict, queries, batches=100, subjects=2000, epochs=8, seed=0):
    """Trains a CostModel for databases like `spec` and the sample `queries`.
    Returns (model, word counts, per-epoch losses)."""
    counts = word_counts(queries)
    X, y = training_set(ctx, spec, queries, counts, batches, subjects)
    model = CostModel(0, seed)
    losses = model.train(X, y, epochs=epochs, batch=4096, lr=3e-3, seed=seed)
    return model, counts, losses


def predicted_cost(model: CostModel, X, lengths, n_queries, n_sample, n_groups):
    """Scan work per subject for the whole batch: predicted hit/extension work
    scaled from the sample to all queries, plus ALPHA per residue per pass."""
    hitwork = np.expm1(np.maximum(model.predict(X), 0.0)).astype(np.float64) * (n_queries / max(n_sample, 1))
    return (hitwork + ALPHA * n_groups * np.asarray(lengths, dtype=np.float64)).astype(np.float32)


def apply_cost(db, cost):
    db_set_cost(db, cost)
