"""GPU cost model: the MFMA MLP kernels against the CPU reference of the same
numerics (tests/cm_reference.py, itself pinned by finite differences), the
feature kernel against a numpy restatement, training, and scheduling by
cost -- which must never change a scan's results."""
import sys

import numpy as np
import pytest

import bpx
import cm_reference as R
from bpx import costsched

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert bpx.available(), "no HIP device visible"
    c = bpx.Context(0)
    yield c
    c.close()


def realistic_inputs(rng, n):
    X = np.zeros((n, R.F), dtype=np.float32)
    X[:, 0] = rng.uniform(3, 10, n)
    comp = rng.dirichlet(np.ones(20), n)
    X[:, 1:21] = comp
    X[:, 21] = rng.uniform(0, 12, n)
    return X


def test_forward_matches_reference(ctx):
    rng = np.random.default_rng(1)
    m = bpx.CostModel(0, 5)
    p = (rng.standard_normal(R.NPARAMS) * 0.15).astype(np.float32)
    m.params = p
    np.testing.assert_array_equal(m.params, p)
    for n in (1, 31, 32, 1000, 4099):
        X = realistic_inputs(rng, n)
        want, _ = R.forward(p, X)
        got = m.predict(X)
        say(f"n={n}: max |diff| {np.abs(got - want).max():.3g}, max |y| {np.abs(want).max():.3g}")
        np.testing.assert_allclose(got, want, rtol=2e-3, atol=2e-3)
    m.close()


def test_gradient_matches_reference(ctx):
    rng = np.random.default_rng(2)
    m = bpx.CostModel(0, 6)
    p = (rng.standard_normal(R.NPARAMS) * 0.15).astype(np.float32)
    m.params = p
    for n in (37, 4096, 5000):
        X = realistic_inputs(rng, n)
        t = rng.uniform(0, 8, n).astype(np.float32)
        loss, g = m.grad(X, t)
        rloss, rg = R.loss_grad(p, X, t)
        assert abs(loss - rloss) <= 1e-3 * max(1.0, rloss), (loss, rloss)
        for name, lo, hi in (("W1", R.OFF_W1, R.OFF_B1), ("b1", R.OFF_B1, R.OFF_W2), ("W2", R.OFF_W2, R.OFF_B2),
                             ("b2", R.OFF_B2, R.OFF_W3), ("W3", R.OFF_W3, R.OFF_B3), ("b3", R.OFF_B3, R.NPARAMS)):
            a, b = g[lo:hi], rg[lo:hi]
            tol = 2e-3 * max(np.abs(b).max(), 1e-6) + 1e-6
            say(f"n={n} {name}: max |diff| {np.abs(a - b).max():.3g} (tol {tol:.3g})")
            assert np.abs(a - b).max() <= tol, name
    m.close()


def test_training_learns_a_known_function(ctx):
    rng = np.random.default_rng(3)
    n = 40000
    X = realistic_inputs(rng, n)
    t = (0.5 * X[:, 0] + 0.3 * X[:, 21] + 4 * X[:, 3] - 2 * X[:, 7]).astype(np.float32)
    m = bpx.CostModel(0, 7)
    losses = m.train(X, t, epochs=10, batch=2048, lr=3e-3, seed=1)
    say(f"losses {losses}")
    assert losses[-1] < 0.05 * losses[0] and losses[-1] < 0.05
    m.close()


def small_skewed_spec(seed, n):
    return dict(seed=seed, num_subjects=n, len_median=350, len_sigma=1.2, len_min=20, len_max=35000,
                homolog_fraction=0.01, identity_min=0.4, identity_max=0.6, indel_min=0.0, indel_max=0.03,
                biased_fraction=0.2, dirichlet_alpha=0.3, motif_min=5, motif_max=30)


def test_features_match_reference(ctx):
    qs = [bpx.synth_query(40, i, L) for i, L in enumerate((60, 300, 900))]
    sdb = bpx.synth_db(small_skewed_spec(41, 700), qs, nthreads=16)
    ddb = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
    counts = bpx.word_counts(qs)
    got = bpx.db_features(ctx, ddb, counts)
    want = R.features(sdb.residues, sdb.offsets, sdb.lengths, counts)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5)
    ddb.close()


def test_cost_order_does_not_change_results(ctx):
    rng = np.random.default_rng(4)
    qs = [bpx.synth_query(42, i, int(L)) for i, L in enumerate(rng.integers(30, 800, 20))]
    sdb = bpx.synth_db(small_skewed_spec(43, 3000), qs, nthreads=16)
    ddb = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
    b = ctx.prepare_batch(qs, group_residues=1200, group_queries=32)

    def run():
        r = bpx.scan_batch(ctx, b, ddb, subject_stats=True).wait()
        recs, subs, qids = r.records(with_queries=True)
        out = (recs.view(np.int32).copy(), subs.copy(), qids.copy(), r.subject_stats().view(np.int32).copy())
        r.close()
        return out

    base = run()
    for cost in (rng.uniform(0, 10, sdb.num_subjects), sdb.lengths.astype(np.float64)[::-1].copy(), None):
        bpx.db_set_cost(ddb, None if cost is None else cost.astype(np.float32))
        got = run()
        for x, y in zip(got, base):
            np.testing.assert_array_equal(x, y)
    b.close()
    ddb.close()


def test_learned_cost_ranks_work_better_than_length(ctx):
    """What the network gains: its predicted work ranks held-out subjects of a
    skewed database closer to their true work than length alone does."""
    lens = bpx.synth_query_lengths(44, 64, "loguniform", 50, 2000)
    qs = [bpx.synth_query(44, i, int(L)) for i, L in enumerate(lens)]
    sample = costsched.sample_queries(qs, 8)
    spec = small_skewed_spec(0, 0)
    model, counts, losses = costsched.fit(ctx, spec, sample, batches=12, subjects=1500, epochs=10)
    say(f"training losses {losses}")
    sdb = bpx.synth_db(small_skewed_spec(99, 4000), sample, nthreads=16)
    ddb = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
    X = bpx.db_features(ctx, ddb, counts)
    truth = costsched.work_labels(ctx, ddb, sample)
    pred = model.predict(X)

    def spearman(a, b):
        ra = np.argsort(np.argsort(a)).astype(np.float64)
        rb = np.argsort(np.argsort(b)).astype(np.float64)
        return float(np.corrcoef(ra, rb)[0, 1])

    s_pred, s_len = spearman(pred, truth), spearman(sdb.lengths, truth)
    say(f"spearman vs true work: model {s_pred:.4f}, length {s_len:.4f}")
    # rank disagreement (1 - rho) at least a third smaller than length's
    assert 1.0 - s_pred < (2.0 / 3.0) * (1.0 - s_len)
    model.close()
    ddb.close()
