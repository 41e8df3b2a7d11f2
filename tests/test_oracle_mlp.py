"""Pin of the MLP oracle: its analytic FP64 gradient equals central finite differences."""
import numpy as np


def test_mlp_oracle_gradient_matches_finite_differences():
    import mlp_oracle as M
    rng = np.random.default_rng(0)
    D, H1, H2, L, NB, B = 5, 7, 6, 9, 3, 11
    W = [rng.standard_normal((H1, D)), rng.standard_normal((H2, H1)), rng.standard_normal((L + NB, H2))]
    b = [rng.standard_normal(H1) * 0.1, rng.standard_normal(H2) * 0.1, rng.standard_normal(L + NB) * 0.1]
    X = rng.standard_normal((B, D))
    yl = rng.integers(0, L, B)
    yb = rng.integers(0, NB, B)
    _, gW, gb = M.forward_backward(W, b, X, yl, yb, L, NB, 0.5, half=False)
    eps = 1e-6
    for l in range(3):
        for idx in [(0, 0), (W[l].shape[0] - 1, W[l].shape[1] - 1), (1, 2)]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][idx] += eps
            Wm[l][idx] -= eps
            fd = (M.loss_only(Wp, b, X, yl, yb, L, NB, 0.5) - M.loss_only(Wm, b, X, yl, yb, L, NB, 0.5)) / (2 * eps)
            assert abs(fd - gW[l][idx]) < 1e-6 * max(1.0, abs(fd))
        for i in [0, len(b[l]) - 1]:
            bp = [x.copy() for x in b]
            bm = [x.copy() for x in b]
            bp[l][i] += eps
            bm[l][i] -= eps
            fd = (M.loss_only(W, bp, X, yl, yb, L, NB, 0.5) - M.loss_only(W, bm, X, yl, yb, L, NB, 0.5)) / (2 * eps)
            assert abs(fd - gb[l][i]) < 1e-6 * max(1.0, abs(fd))
