"""CPU restatement of the schedule selector — TEST INFRASTRUCTURE ONLY.

The checker of taco_amd's selector kernels (taco_amd/csrc/selector.hip):
  csr_stats / features_from_stats  the integer sparsity statistics and the
      32 features, restated in numpy from their definitions;
  forward / loss_and_grad  the MLP 32 -> 128 -> 128 -> 25 in float64 with
      its masked-softmax cross-entropy + weighted log-time loss and the
      hand-derived backward pass; check_gradients compares that backward
      pass with central finite differences.
There is no reference counterpart (taco has no learned components).
"""
import numpy as np

N_IN, N_HID, N_OUT, N_CLASSES, N_SPMV = 32, 128, 25, 24, 13
TIME_WEIGHT = 0.1
SAMPLE_ROWS = 4096


def csr_stats(pos, crd, m, n):
    pos = np.asarray(pos, dtype=np.int64)
    crd = np.asarray(crd, dtype=np.int64)
    s = np.zeros(30, dtype=np.int64)
    lens = np.diff(pos)
    s[0], s[1], s[2] = m, n, pos[-1] if m else 0
    s[3] = lens.sum()
    s[4] = (lens * lens).sum()
    s[5] = lens.max() if m else 0
    s[6] = (lens == 0).sum()
    bins = np.zeros(m, dtype=np.int64)
    nz = lens > 0
    bins[nz] = np.minimum(np.floor(np.log2(lens[nz])).astype(np.int64) + 1, 15)
    s[7:23] = np.bincount(bins, minlength=16)[:16]
    s[23] = lens[lens >= 64].sum()
    s[24] = lens[lens >= 1024].sum()
    if m and n:
        S = min(m, SAMPLE_ROWS)
        rows = np.arange(S, dtype=np.int64) * m // S
        s[25] = S
        for r in rows:
            c = crd[pos[r]:pos[r + 1]]
            diag = r * n // m
            s[26] += c.size
            s[27] += np.count_nonzero(np.abs(c - diag) <= 64)
            s[28] += np.count_nonzero(c // 8 == r // 8)
            s[29] += (np.abs(c * m - r * n) * 4096 // (n * m)).sum()
    return s


def features_from_stats(s, op, K):
    m, n, nnz = float(s[0]), float(s[1]), float(s[2])
    mean = s[3] / m if m else 0.0
    var = max(0.0, s[4] / m - mean * mean) if m else 0.0
    mx = float(s[5])
    sampled = float(s[26])
    f = np.zeros(N_IN, dtype=np.float64)
    f[0] = op
    f[1] = np.log2(K if K > 0 else 1) if op == 1 else 0.0
    f[2] = np.log2(m + 1)
    f[3] = np.log2(n + 1)
    f[4] = np.log2(nnz + 1)
    f[5] = np.log2(1 + mean)
    f[6] = np.sqrt(var) / mean if mean > 0 else 0.0
    f[7] = np.log2(1 + mx)
    f[8] = np.log2(1 + (mx / mean if mean > 0 else 0.0))
    f[9] = s[6] / m if m else 0.0
    f[10:26] = s[7:23] / m if m else 0.0
    f[26] = s[23] / nnz if nnz else 0.0
    f[27] = s[24] / nnz if nnz else 0.0
    f[28] = s[29] / (4096.0 * sampled) if sampled else 0.0
    f[29] = s[27] / sampled if sampled else 0.0
    f[30] = s[28] / sampled if sampled else 0.0
    f[31] = np.log2(1 + n * (4.0 * (K if K > 0 else 1) if op == 1 else 8.0))
    return f.astype(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    shapes = [(N_HID, N_IN), (N_HID,), (N_HID, N_HID), (N_HID,), (N_OUT, N_HID), (N_OUT,)]
    out, at = [], 0
    for sh in shapes:
        k = int(np.prod(sh))
        out.append(p[at:at + k].reshape(sh))
        at += k
    assert at == p.size
    return out


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return X, h1, h2, h2 @ W3.T + b3


def loss_and_grad(p, X, cls, logt):
    """Mean loss over the batch and its gradient (flat, float64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X, h1, h2, Z = forward(p, X)
    B = X.shape[0]
    dZ = np.zeros_like(Z)
    loss = 0.0
    for b in range(B):
        c = int(cls[b])
        lo, hi = (0, N_SPMV) if c < N_SPMV else (N_SPMV, N_CLASSES)
        z = Z[b, lo:hi]
        mx = z.max()
        lse = mx + np.log(np.exp(z - mx).sum())
        pr = np.exp(z - lse)
        pr[c - lo] -= 1.0
        dZ[b, lo:hi] = pr / B
        d = Z[b, N_CLASSES] - logt[b]
        dZ[b, N_CLASSES] = 2 * TIME_WEIGHT * d / B
        loss += (lse - Z[b, c]) + TIME_WEIGHT * d * d
    loss /= B
    dW3 = dZ.T @ h2
    db3 = dZ.sum(axis=0)
    d2 = (dZ @ W3) * (h2 > 0)
    dW2 = d2.T @ h1
    db2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * (h1 > 0)
    dW1 = d1.T @ X
    db1 = d1.sum(axis=0)
    return loss, np.concatenate([g.ravel() for g in (dW1, db1, dW2, db2, dW3, db3)])


def check_gradients(p, X, cls, logt, idx, h=1e-6):
    """Max difference between loss_and_grad and central finite differences
    at parameter indices idx, relative to |fd| + |grad| floored at 1e-4
    (the differences' own rounding error is about 1e-16 * loss / h)."""
    p = np.asarray(p, dtype=np.float64).copy()
    _, g = loss_and_grad(p, X, cls, logt)
    worst = 0.0
    for i in idx:
        old = p[i]
        p[i] = old + h
        lp, _ = loss_and_grad(p, X, cls, logt)
        p[i] = old - h
        lm, _ = loss_and_grad(p, X, cls, logt)
        p[i] = old
        fd = (lp - lm) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(1e-4, abs(fd) + abs(g[i])))
    return worst
