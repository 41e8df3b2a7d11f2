"""NumPy FP64 oracle of the learned entry-point MLP (test infrastructure only).

Architecture: D -> H1 -> H2 -> (L + NB), ReLU, two softmax cross-entropy heads
(landmarks [0, L) and beam-size buckets [L, L+NB)), loss = mean CE_landmark +
bucket_weight * mean CE_bucket.  `half=True` rounds exactly the tensors the GPU
kernels hold in FP16 (inputs, weights used as operands, activations, backward
operands) so the GPU gradients can be compared to this oracle; `half=False` is
the plain FP64 model whose analytic gradient is pinned by finite differences.
"""
import numpy as np


def _h(a, half):
    return a.astype(np.float16).astype(np.float64) if half else a.astype(np.float64)


def forward_backward(W, b, X, yl, yb, L, NB, bucket_weight=1.0, half=True):
    """W, b: lists of 3 arrays ([out x in], [out]).  Returns (loss, gW, gb)."""
    B = X.shape[0]
    X16 = _h(X, half)
    Wh = [_h(w, half) for w in W]
    bd = [x.astype(np.float64) for x in b]
    P1 = X16 @ Wh[0].T + bd[0]
    A1 = _h(np.maximum(P1, 0), half)
    P2 = A1 @ Wh[1].T + bd[1]
    A2 = _h(np.maximum(P2, 0), half)
    Z = A2 @ Wh[2].T + bd[2]
    dZ = np.zeros_like(Z)
    loss = 0.0
    for s, e, y, wgt in ((0, L, yl, 1.0), (L, L + NB, yb, bucket_weight)):
        z = Z[:, s:e]
        mx = z.max(1, keepdims=True)
        ez = np.exp(z - mx)
        se = ez.sum(1, keepdims=True)
        p = ez / se
        loss += wgt * float(np.mean(np.log(se[:, 0]) + mx[:, 0] - z[np.arange(B), y]))
        g = p.copy()
        g[np.arange(B), y] -= 1.0
        dZ[:, s:e] = wgt * g / B
    dZ16 = _h(dZ, half)
    gW3 = dZ16.T @ A2
    gb3 = dZ.sum(0)
    dA2 = dZ16 @ Wh[2]
    dH2 = dA2 * (P2 > 0)
    dH2_16 = _h(dH2, half)
    gb2 = dH2.sum(0)
    gW2 = dH2_16.T @ A1
    dA1 = dH2_16 @ Wh[1]
    dH1 = dA1 * (P1 > 0)
    dH1_16 = _h(dH1, half)
    gb1 = dH1.sum(0)
    gW1 = dH1_16.T @ X16
    return loss, [gW1, gW2, gW3], [gb1, gb2, gb3]


def loss_only(W, b, X, yl, yb, L, NB, bucket_weight=1.0):
    return forward_backward(W, b, X, yl, yb, L, NB, bucket_weight, half=False)[0]
