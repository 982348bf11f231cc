"""CPU reference of the cost-model MLP (32-64-64-1, ReLU) with the library's
numerics: bf16 (round to nearest even) weights, inputs, activations and
back-propagated errors at the points where the GPU kernels feed MFMA, fp32
accumulation and fp32 biases.  With rnd=False every rounding is the identity,
which is what the finite-difference check of this reference uses."""
import numpy as np

F, H, O = 32, 64, 16
OFF_W1, OFF_B1 = 0, F * H
OFF_W2, OFF_B2 = OFF_B1 + H, OFF_B1 + H + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * O
NPARAMS = OFF_B3 + O


def bf16(x, rnd=True):
    x = np.asarray(x, dtype=np.float32)
    if not rnd:
        return x
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(H, O), p[OFF_B3:NPARAMS])


def forward(p, X, rnd=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Xb = bf16(X, rnd)
    H1 = bf16(np.maximum(Xb @ bf16(W1, rnd) + b1, 0), rnd)
    H2 = bf16(np.maximum(H1 @ bf16(W2, rnd) + b2, 0), rnd)
    y = H2 @ bf16(W3, rnd)[:, 0] + b3[0]
    return y.astype(np.float32), (Xb, H1, H2)


def loss_grad(p, X, t, rnd=True):
    """Mean squared error and its gradient w.r.t. the parameter vector."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = len(t)
    y, (Xb, H1, H2) = forward(p, X, rnd)
    e = (y - t).astype(np.float32)
    dy = (np.float32(2.0 / n) * e).astype(np.float32)
    g = np.zeros(NPARAMS, dtype=np.float32)
    dz3 = bf16(dy, rnd)
    g[OFF_B3] = dy.sum(dtype=np.float32)
    gW3 = np.zeros((H, O), dtype=np.float32)
    gW3[:, 0] = H2.T @ dz3
    g[OFF_W3:OFF_B3] = gW3.ravel()
    dH2 = np.outer(dz3, bf16(W3, rnd)[:, 0]).astype(np.float32)
    dZ2 = dH2 * (H2 > 0)
    g[OFF_B2:OFF_W3] = dZ2.sum(axis=0, dtype=np.float32)
    dZ2b = bf16(dZ2, rnd)
    g[OFF_W2:OFF_B2] = (H1.T @ dZ2b).ravel()
    dH1 = dZ2b @ bf16(W2, rnd).T
    dZ1 = dH1 * (H1 > 0)
    g[OFF_B1:OFF_W2] = dZ1.sum(axis=0, dtype=np.float32)
    g[OFF_W1:OFF_B1] = (Xb.T @ bf16(dZ1, rnd)).ravel()
    return float(np.mean(e.astype(np.float64) ** 2)), g


def features(residues, offsets, lengths, counts, word=3):
    """The per-subject feature rows the GPU feature kernel computes."""
    std20 = [1, 16, 13, 4, 3, 15, 5, 7, 8, 9, 11, 10, 12, 6, 14, 17, 18, 20, 22, 19]
    out = np.zeros((len(lengths), F), dtype=np.float32)
    for i, L in enumerate(lengths):
        s = residues[int(offsets[i]):int(offsets[i]) + int(L)].astype(np.int64)
        out[i, 0] = np.log1p(np.float32(L))
        if L:
            hist = np.bincount(s & 31, minlength=32)
            out[i, 1:21] = hist[std20] / np.float32(L)
        if L >= word:
            w = np.zeros(L - word + 1, dtype=np.int64)
            for k in range(word):
                w = (w << 5) | s[k:L - word + 1 + k]
            out[i, 21] = np.log1p(np.float32(counts[w].astype(np.float64).sum()))
    return out
