"""The cost scheduler's host arithmetic on the CPU: which queries form the
fixed sample, and how a model's per-subject prediction becomes the batch's
predicted work (sample -> whole batch scaling, the per-residue pass term,
negative predictions clamped).  A stub stands in for the GPU model."""
import numpy as np
import pytest

from bpx import costsched


class StubModel:
    """predict(X) returns the first feature column: the log1p work."""

    def predict(self, X):
        return np.asarray(X, dtype=np.float32)[:, 0]


def test_sample_queries_takes_every_kth_and_at_most_n():
    qs = [np.full(i + 1, 1, np.uint8) for i in range(100)]
    s = costsched.sample_queries(qs, 16)
    assert len(s) == 16
    assert [len(q) for q in s] == [1 + 6 * i for i in range(16)]  # step 100 // 16 = 6
    few = qs[:5]
    assert [len(q) for q in costsched.sample_queries(few, 16)] == [1, 2, 3, 4, 5]


def test_predicted_cost_scales_the_sample_and_adds_the_pass_term():
    work = np.array([0.0, 10.0, 250.0, 3.0])
    X = np.log1p(work)[:, None].astype(np.float32)
    lengths = np.array([100, 200, 300, 400])
    got = costsched.predicted_cost(StubModel(), X, lengths, n_queries=1024, n_sample=16, n_groups=375)
    want = work * (1024 / 16) + costsched.ALPHA * 375 * lengths
    np.testing.assert_allclose(got, want, rtol=1e-5)
    assert got.dtype == np.float32


def test_predicted_cost_clamps_negative_predictions():
    X = np.array([[-3.0], [-0.5]], dtype=np.float32)  # a regression can undershoot log1p(0) = 0
    got = costsched.predicted_cost(StubModel(), X, [10, 20], n_queries=8, n_sample=8, n_groups=2)
    np.testing.assert_allclose(got, costsched.ALPHA * 2 * np.array([10, 20]), rtol=1e-6)


def test_predicted_cost_with_an_empty_sample_does_not_divide_by_zero():
    X = np.log1p(np.array([[4.0]], dtype=np.float32))
    got = costsched.predicted_cost(StubModel(), X, [50], n_queries=3, n_sample=0, n_groups=1)
    assert got[0] == pytest.approx(4.0 * 3 + costsched.ALPHA * 50, rel=1e-5)
