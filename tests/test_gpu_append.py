"""append_items! on an exhaustive index (ssg_append_db): appending to an uploaded index
gives the same results, bit for bit, as uploading every object at once, across capacity
doublings, for the exact and the precision forms; a graph handle refuses appends."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ss():
    import similaritysearch as ss
    return ss


def test_appends_equal_one_upload(oracle):
    ss = _ss()
    rng = np.random.default_rng(41)
    X = rng.random((9000, 24), dtype=np.float32)
    Q = ss.MatrixDatabase.from_rows(rng.random((400, 24), dtype=np.float32))
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(np.zeros((0, 24), np.float32)))
    E.append_items(None, ss.MatrixDatabase.from_rows(X[:10]))       # before the upload
    E._handle()
    cuts = [10, 11, 75, 64 * 7 + 3, 3000, 3001, 9000]
    for a, b in zip(cuts, cuts[1:]):                                  # after it, across doublings
        E.append_items(None, ss.MatrixDatabase.from_rows(X[a:b]))
    E.push_item(None, X[0])
    assert len(E) == 9001
    full = np.concatenate([X, X[:1]])
    ids, dists = ss.searchbatch(E, ss.GenericContext(), Q, 10)
    gi, gd = oracle.exhaustive("SqL2", full, Q.rows, 10)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(dists.T, gd)
    F = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(full))
    for fmt in ("bf16", "fp8"):
        a = ss.exhaustive_candidates(E, 10, 64, Q, precision=fmt)
        b = ss.exhaustive_candidates(F, 10, 64, Q, precision=fmt)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    E.append_items(None, ss.MatrixDatabase.from_rows(X[:500]))        # after the encoded copies
    a = ss.exhaustive_candidates(E, 10, 64, Q, precision="fp8")
    F2 = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(np.concatenate([full, X[:500]])))
    b = ss.exhaustive_candidates(F2, 10, 64, Q, precision="fp8")
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_graph_handle_refuses_appends():
    ss = _ss()
    from similaritysearch import _lib
    X = np.random.default_rng(42).random((2000, 8), dtype=np.float32)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), seed=2)
    rc = _lib.load().ssg_append_db(G._handle(), X[:1].ctypes.data, 1)
    assert rc != 0 and b"graph" in _lib.load().ssg_last_error(None)
