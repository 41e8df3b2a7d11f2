"""Host-side pieces of the boundary (no GPU): the KnnSorted / KnnHeap queues the
queue-vector searchbatch! fills (src/pqueue/knnsorted.jl, knnheap.jl, heap.jl), the
AdjList / AdjDict -> StaticAdjList conversion (src/adj/*.jl), and the Cosine row
normalisation against the oracle's."""
import numpy as np
import pytest


def _ss():
    import similaritysearch as ss
    return ss


@pytest.mark.parametrize("k", [1, 4, 10, 33])
def test_knnsorted_matches_oracle_queue(oracle, k):
    """Random pushes and pops against or_knn_run (pinned by sort-based gold)."""
    from similaritysearch.pqueue import KnnSorted, knnqueue
    rng = np.random.default_rng(k)
    nops = 400
    ops = (rng.random(nops) < 0.15).astype(np.int8)
    ids = np.arange(1, nops + 1, dtype=np.uint32)
    d = rng.integers(0, 30, nops).astype(np.float32)   # many ties
    q = knnqueue(KnnSorted, k)
    acc = []
    for t in range(nops):
        if ops[t] == 1:
            if len(q) == 0:
                ops[t] = 0
            else:
                q.pop_min()
                acc.append(1)
                continue
        acc.append(int(q.push_item(ids[t], d[t])))
    oi, od, oacc = oracle.knn_run(k, ops, ids, d)
    np.testing.assert_array_equal(q.id_view(), oi)
    np.testing.assert_array_equal(q.dist_view(), od)
    np.testing.assert_array_equal(np.array(acc, bool), oacc)


@pytest.mark.parametrize("k", [1, 2, 7, 16])
def test_knnheap_keeps_the_k_nearest(k):
    """Distinct distances: the heap holds the k nearest with the farthest at the root, the
    nearest tracked apart, and sortitems returns them ascending; a full heap rejects a
    push at its maximum."""
    from similaritysearch.pqueue import KnnHeap, knnqueue
    rng = np.random.default_rng(10 + k)
    d = rng.permutation(200).astype(np.float32)
    q = knnqueue(KnnHeap, k)
    for i, x in enumerate(d):
        q.push_item(i + 1, x)
        n = len(q)
        D = q.dists[:n]
        for p in range(1, n + 1):               # max-heap property
            for c in (2 * p, 2 * p + 1):
                if c <= n:
                    assert D[p - 1] >= D[c - 1]
    assert q.nearest() == (int(np.argmin(d)) + 1, float(d.min()))
    assert not q.push_item(999, q.maximum())
    ids, ds = q.sortitems()
    np.testing.assert_array_equal(ids, np.argsort(d)[:k] + 1)
    np.testing.assert_array_equal(ds, np.sort(d)[:k])


def test_adjacency_conversions():
    ss = _ss()
    S = ss.static_adjlist(ss.AdjList([[2, 3], [1], None, [1, 2]]))
    np.testing.assert_array_equal(S.offset, [2, 3, 3, 5])
    np.testing.assert_array_equal(S.end_point, [2, 3, 1, 1, 2])
    assert S.neighbors_length(3) == 0
    S = ss.static_adjlist(ss.AdjDict({1: [2], 4: [3, 1]}, 4))
    np.testing.assert_array_equal(S.offset, [1, 1, 1, 3])
    np.testing.assert_array_equal(S.neighbors(4), [3, 1])
    with pytest.raises(TypeError):
        ss.static_adjlist([[1], [2]])


def test_cosine_rows_match_oracle(oracle):
    ss = _ss()
    rng = np.random.default_rng(52)
    R = (rng.standard_normal((500, 40)) * rng.uniform(0.01, 100, (500, 1))).astype(np.float32)
    R[7] = 0
    np.testing.assert_array_equal(ss.unit_rows(R), oracle.unit_rows(R))
