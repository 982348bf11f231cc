"""The CPU reference of the cost-model MLP is itself pinned: without bf16
rounding its analytic gradient equals central finite differences (float64
re-evaluation), and its rounding helper is round-to-nearest-even."""
import numpy as np

import cm_reference as R


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, 3.0e-38], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3*2^-8 ties between 1+2^-7 and 1+2^-6: even (1 + 2^-6) wins
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0], dtype=np.float32)
    np.testing.assert_array_equal(R.bf16(x)[:6], want)
    assert R.bf16(x)[6] > 0


def test_reference_gradient_equals_finite_differences():
    rng = np.random.default_rng(0)
    p = (rng.standard_normal(R.NPARAMS) * 0.2).astype(np.float32)
    X = rng.standard_normal((40, R.F)).astype(np.float32)
    t = rng.standard_normal(40).astype(np.float32)
    _, g = R.loss_grad(p, X, t, rnd=False)

    def loss64(q):
        W1, b1, W2, b2, W3, b3 = [a.astype(np.float64) for a in R.unpack(q)]
        h1 = np.maximum(X @ W1 + b1, 0)
        h2 = np.maximum(h1 @ W2 + b2, 0)
        y = h2 @ W3[:, 0] + b3[0]
        return np.mean((y - t) ** 2)

    idx = np.concatenate([rng.choice(R.NPARAMS, 60, replace=False),
                          [R.OFF_B1, R.OFF_B2, R.OFF_B3, R.OFF_W3, R.OFF_W3 + 16]])
    for i in idx:
        h = 1e-3
        qp, qm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        qp[i] += h
        qm[i] -= h
        fd = (loss64(qp) - loss64(qm)) / (2 * h)
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(fd)), (int(i), fd, g[i])
    # parameters that cannot affect the output have zero gradient
    W3cols = np.arange(R.OFF_W3, R.OFF_B3).reshape(R.H, R.O)[:, 1:].ravel()
    assert np.all(g[W3cols] == 0) and np.all(g[R.OFF_B3 + 1:] == 0)
