"""rerank! on long columns (65..1024 rows; src/rerank.jl:22-41, 77-86): the GPU result equals
the oracle's stable re-sort of the recomputed canonical distances, bit for bit, with columns
that end early (0) and tied distances."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ss():
    import similaritysearch as ss
    return ss


@pytest.mark.parametrize("metric,k,dim", [("SqL2", 100, 24), ("L2", 300, 96), ("NormCosine", 1024, 16),
                                          ("SqL2", 65, 256)])
def test_rerank_long_columns(oracle, metric, k, dim):
    ss = _ss()
    rng = np.random.default_rng(k + dim)
    X = rng.random((3000, dim), dtype=np.float32)
    if metric == "NormCosine":
        X /= np.linalg.norm(X, axis=1, keepdims=True)
    X[10:20] = X[0]                                           # tied distances
    Q = X[rng.integers(0, 3000, 50)] + np.float32(0.01)
    ids = np.stack([rng.permutation(3000)[:k] + 1 for _ in range(50)]).astype(np.uint32)
    ids[::3, 0:10] = np.arange(11, 21, dtype=np.uint32)       # columns holding the ties
    ids[1, k // 2:] = 0                                       # columns ending early
    ids[4, 1:] = 0
    ids[7, :] = 0
    dists = np.zeros(ids.shape, np.float32)
    dist = {"SqL2": ss.SqL2, "L2": ss.L2, "NormCosine": ss.NormCosine}[metric]
    I = np.asfortranarray(ids.T.copy())
    D = np.asfortranarray(dists.T.copy())
    ss.rerank_(dist, ss.MatrixDatabase.from_rows(X), ss.MatrixDatabase.from_rows(Q), I, D)
    gi, gd = oracle.rerank(metric, X, Q, ids, dists)
    np.testing.assert_array_equal(I.T, gi)
    np.testing.assert_array_equal(D.T, gd)
    assert (I.T[7] == 0).all()
