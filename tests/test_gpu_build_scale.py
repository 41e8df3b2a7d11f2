"""GPU index! at a size where every visited slot wraps its epoch (> 255 x 128 inserts):
two builds are identical and equal the oracle's graph, hints and tuned BeamSearch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_build_60k_deterministic_and_exact(oracle):
    import similaritysearch as ss
    from similaritysearch import datasets
    X = datasets.gmm(60000, 64, seed=4)
    G1 = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=128, seed=1)
    G2 = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=128, seed=1)
    np.testing.assert_array_equal(G1.adj.end_point, G2.adj.end_point)
    g = oracle.build_graph("SqL2", X, parallel_block=128, seed=1)
    np.testing.assert_array_equal(G1.adj.offset, g["offs"][1:])
    np.testing.assert_array_equal(G1.adj.end_point, g["adj"])
    np.testing.assert_array_equal(G1.hints, g["hints"])
    assert (G1.algo.bsize, G1.algo.maxvisits) == (g["bsize"], g["maxvisits"])
