"""GPU parity of the batch consumers of the search path: closestpair /
bichromatic_closestpair (src/bichromatic/*.jl), neardup (src/neardup.jl) and hsp_queries
(src/hsp.jl).  Every distance they use comes from the GPU (searchbatch! or rerank!
through the C ABI); the results are compared with the oracle's restatements
(oracle/pyoracle.py, pinned by brute force and invariants in tests/test_oracle.py),
bit for bit: ids, distances, maps and evaluation counts.
"""
import numpy as np
import pytest


def _ss():
    import similaritysearch as ss
    return ss


def _data(n, dim, seed, metric="SqL2"):
    X = np.random.default_rng(seed).random((n, dim), dtype=np.float32)
    if metric == "NormCosine":
        X /= np.linalg.norm(X, axis=1, keepdims=True)
    return X


def _db(X):
    return _ss().MatrixDatabase.from_rows(X)         # (n, dim) rows = the reference's (dim, n)


def _dist(ss, metric):
    return {"SqL2": ss.SqL2(), "L2": ss.L2(), "NormCosine": ss.NormCosine()}[metric]


def test_positions_host_logic():
    """The host-side map from rerank!'s sorted columns back to object positions."""
    from similaritysearch.consumers import _positions
    rows = np.array([[5, 9, 2, 0], [7, 3, 1, 4]], np.uint32)
    found = np.array([[2, 5, 9, 0], [4, 1, 7, 3]], np.uint32)
    np.testing.assert_array_equal(_positions(rows, found), [[2, 0, 1, -1], [3, 2, 0, 1]])


@pytest.mark.gpu
@pytest.mark.parametrize("metric", ["SqL2", "L2", "NormCosine"])
def test_closestpair_exhaustive(oracle, metric):
    ss = _ss()
    X = _data(2000, 12, 80, metric)
    E = ss.ExhaustiveSearch(_dist(ss, metric), _db(X))
    got = ss.closestpair(E, ss.GenericContext())
    assert got == oracle.bichromatic_closestpair(metric, X, X, 8, samedata=True)
    X[1500] = X[37]                                   # a planted duplicate pair
    E = ss.ExhaustiveSearch(_dist(ss, metric), _db(X))
    got = ss.closestpair(E, ss.GenericContext(), min_k=2)
    assert got == oracle.bichromatic_closestpair(metric, X, X, 2, samedata=True)
    assert got[:2] == (1501, 38)


@pytest.mark.gpu
def test_bichromatic_closestpair_exhaustive(oracle):
    ss = _ss()
    A = _data(3000, 8, 81)
    B = _data(700, 8, 82)
    E = ss.ExhaustiveSearch(ss.SqL2(), _db(A))
    for min_k in (1, 8):
        got = ss.bichromatic_closestpair(E, ss.GenericContext(), _db(B), min_k=min_k)
        assert got == oracle.bichromatic_closestpair("SqL2", A, B, min_k)
    with pytest.raises(ValueError):
        ss.closestpair(E, ss.GenericContext(), min_k=1)


@pytest.mark.gpu
def test_closestpair_searchgraph(oracle):
    """Over a SearchGraph: each object's search starts from a random neighbor with the
    object marked visited (bichromaticclosestpair.jl:107-113); the same hints and marks in
    the oracle's beam search give the same rows and the same pair.  A planted duplicate
    pair, far below every other distance, is found (from either side)."""
    ss = _ss()
    from similaritysearch.consumers import closestpair_hints
    X = _data(3000, 8, 83)
    X[2500] = X[100] + np.float32(1e-4)
    G = ss.SearchGraph.build(ss.SqL2(), _db(X), seed=3)
    ctx = ss.SearchGraphContext()
    got = ss.closestpair(G, ctx, seed=5)
    H = closestpair_hints(G, 5)
    qh = np.stack([h[:1] for h in H]).astype(np.uint32)
    offs = np.concatenate([[0], G.adj.offset]).astype(np.int64)
    ids, d, _, _ = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, X, 8, G.algo.bsize,
                                     G.algo.delta, G.algo.maxvisits, beamcap=ctx.beamcap, qhints=qh,
                                     qh_count=np.ones(3000, np.int32),
                                     qvisit=np.arange(1, 3001, dtype=np.uint32), nthreads=8)
    assert not (ids == np.arange(1, 3001, dtype=np.uint32)[:, None]).any()   # never itself
    assert got == oracle.closest_of_rows(ids, d, True)
    E = ss.ExhaustiveSearch(ss.SqL2(), _db(X))
    exact = ss.closestpair(E, ss.GenericContext())
    # the approximate search may find the pair from either side
    assert exact[:2] == (2501, 101) and {got[0], got[1]} == {101, 2501} and got[2] == exact[2]


@pytest.mark.gpu
def test_searchbatch_knns_visited_marks(oracle):
    """Per-query visited marks: the marked object never enters the result or the beam,
    and the search equals the oracle's with the same marks."""
    ss = _ss()
    X = _data(5000, 16, 95)
    Q = X[:400] + np.float32(1e-3)
    G = ss.SearchGraph.build(ss.SqL2(), _db(X), seed=6)
    ctx = ss.SearchGraphContext()
    vis = np.arange(1, 401, dtype=np.uint32)
    vis[::5] = 0                                                  # some queries mark nothing
    res = ss.searchbatch_knns(G, ctx, _db(Q), np.full(400, 10, np.int32), visited=vis)
    ids = np.stack([r[0] for r in res])
    dists = np.stack([r[1] for r in res])
    offs = np.concatenate([[0], G.adj.offset]).astype(np.int64)
    gi, gd, gcd, _ = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 10, G.algo.bsize,
                                       G.algo.delta, G.algo.maxvisits, beamcap=ctx.beamcap,
                                       qvisit=vis, nthreads=8)
    np.testing.assert_array_equal(ids, gi)
    np.testing.assert_array_equal(dists, gd)
    np.testing.assert_array_equal(ctx.costdists, gcd)
    marked = vis != 0
    assert not (ids[marked] == vis[marked][:, None]).any()
    assert (ids[~marked][:, 0] == np.nonzero(~marked)[0] + 1).mean() > 0.9   # unmarked find their source


@pytest.mark.gpu
@pytest.mark.parametrize("metric,eps", [("SqL2", 0.02), ("L2", 0.15), ("NormCosine", 0.002)])
@pytest.mark.parametrize("blocksize,filterblocks", [(256, True), (100, True), (128, False)])
def test_neardup_matches_oracle(oracle, metric, eps, blocksize, filterblocks):
    ss = _ss()
    X = _data(1500, 3, 84, metric)
    X[700:760] = X[10:70] + np.float32(1e-5)          # near duplicates across blocks
    R = ss.neardup(_dist(ss, metric), _db(X), eps, blocksize=blocksize, filterblocks=filterblocks)
    M, L, D, centers, cost = oracle.neardup(metric, X, eps, blocksize=blocksize,
                                            filterblocks=filterblocks)
    np.testing.assert_array_equal(R["map"], M)
    np.testing.assert_array_equal(R["nn"], L)
    np.testing.assert_array_equal(R["dist"], D)
    np.testing.assert_array_equal(R["centers"], centers)
    assert R["costdists"] == cost
    assert len(R["idx"]) == len(centers) < 1500


@pytest.mark.gpu
def test_neardup_all_duplicates_and_single_block(oracle):
    ss = _ss()
    X = np.tile(_data(5, 6, 85), (40, 1))             # 5 distinct rows, 40 copies each
    R = ss.neardup(ss.SqL2(), _db(X), 0.0, blocksize=64)
    M, L, D, centers, cost = oracle.neardup("SqL2", X, 0.0, blocksize=64)
    np.testing.assert_array_equal(R["centers"], centers)
    np.testing.assert_array_equal(R["nn"], L)
    assert list(centers) == [1, 2, 3, 4, 5]
    R = ss.neardup(ss.SqL2(), _db(X[:50]), 1e9, blocksize=256)
    assert list(R["centers"]) == [1] and (R["nn"] == 1).all()
    # eps = +Inf over several blocks: one center, every later block is covered by the net
    # search, and each searched block charges the net's size per object
    R = ss.neardup(ss.SqL2(), _db(X), np.inf, blocksize=64)
    M, L, D, centers, cost = oracle.neardup("SqL2", X, np.inf, blocksize=64)
    assert list(R["centers"]) == [1] and (R["nn"] == 1).all()
    np.testing.assert_array_equal(R["dist"], D)
    assert R["costdists"] == cost == 63 + (200 - 64) * 1


@pytest.mark.gpu
@pytest.mark.parametrize("metric", ["SqL2", "L2", "NormCosine"])
def test_hsp_queries_matches_oracle(oracle, metric):
    ss = _ss()
    X = _data(4000, 4, 86, metric)
    Q = _data(300, 4, 87, metric)
    E = ss.ExhaustiveSearch(_dist(ss, metric), _db(X))
    ids, dists = ss.searchbatch(E, ss.GenericContext(), _db(Q), 32)
    hi, hd, hsp = ss.hsp_queries(_dist(ss, metric), _db(X), _db(Q), ids, dists)
    oi, od = oracle.hsp_queries(metric, X, Q, ids.T, dists.T)
    np.testing.assert_array_equal(hi.T, oi)
    np.testing.assert_array_equal(hd.T, od)
    assert [len(h[0]) for h in hsp] == list((oi != 0).sum(1))
    assert 1 < np.mean((oi != 0).sum(1)) < 32


@pytest.mark.gpu
def test_hsp_queries_short_columns(oracle):
    """Candidate columns that end early (k above the database size) and the database
    itself as the query set."""
    ss = _ss()
    X = _data(20, 3, 88)
    E = ss.ExhaustiveSearch(ss.SqL2(), _db(X))
    ids, dists = ss.searchbatch(E, ss.GenericContext(), _db(X), 24)
    assert (ids[20:] == 0).all()
    hi, hd, _ = ss.hsp_queries(ss.SqL2(), _db(X), _db(X), ids, dists)
    oi, od = oracle.hsp_queries("SqL2", X, X, ids.T, dists.T)
    np.testing.assert_array_equal(hi.T, oi)
    np.testing.assert_array_equal(hd.T, od)
    assert (oi[:, 0] == np.arange(1, 21)).all()       # each object keeps itself first


@pytest.mark.gpu
@pytest.mark.parametrize("k", [15, 100])
def test_pairwise_matches_dist_batch(oracle, k):
    ss = _ss()
    X = _data(5000, 7, 89)
    ids = np.random.default_rng(90).permutation(5000)[:20 * k].reshape(20, k).astype(np.uint32) + 1
    ids[3, 9:] = 0
    P = ss.pairwise(ss.SqL2(), _db(X), ids)
    for r in range(20):
        v = ids[r][ids[r] > 0]
        ref = oracle.dist_batch("SqL2", X[v - 1], X[v - 1])
        np.testing.assert_array_equal(P[r, :len(v), :len(v)], ref)
        assert np.isinf(P[r, len(v):]).all() and np.isinf(P[r, :, len(v):]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("n,dim,k", [(3000, 256, 64), (2000, 100, 40), (500, 16, 1), (3000, 96, 100),
                                     (3000, 96, 256), (4000, 256, 160)])
def test_hsp_queries_shapes_and_unsorted_columns(oracle, n, dim, k):
    """Wide rows with k = 64 (one wave per block), ragged dims, k = 1, columns of more than
    64 rows (two and four queue registers per lane, LDS beyond 64 KiB at k = 160, dim 256),
    and columns given out of distance order (the kept queue sorts them, stable on ties)."""
    ss = _ss()
    X = _data(n, dim, 91)
    Q = _data(64, dim, 92)
    ids, dists = (np.ascontiguousarray(a.T) for a in oracle.exhaustive("SqL2", X, Q, k))
    rng = np.random.default_rng(93)
    perm = rng.permutation(k)
    ids, dists = np.ascontiguousarray(ids[perm]), np.ascontiguousarray(dists[perm])
    if k >= 3:
        dists[1:3] = dists[0]                         # ties among the given distances
    hi, hd, _ = ss.hsp_queries(ss.SqL2(), _db(X), _db(Q), ids, dists)
    oi, od = oracle.hsp_queries("SqL2", X, Q, ids.T, dists.T)
    np.testing.assert_array_equal(hi.T, oi)
    np.testing.assert_array_equal(hd.T, od)
    with pytest.raises(Exception, match="256"):
        ss.hsp_queries(ss.SqL2(), _db(X), _db(Q[:2]), np.ones((257, 2), np.uint32),
                       np.zeros((257, 2), np.float32))


@pytest.mark.gpu
def test_hsp_queries_hand_case():
    """Points on a line, query at 0 (the oracle's hand case): equal distance to a kept
    candidate keeps (strict <); the kept come back sorted, stable on ties."""
    ss = _ss()
    X = np.zeros((4, 3), np.float32)
    X[:, 0] = [2.0, 1.0, 3.0, -1.0]
    ids = np.array([[1], [2], [3], [4]], np.uint32)                  # (k, n) = (4, 1)
    d = np.array([[4.0], [1.0], [9.0], [1.0]], np.float32)
    hi, hd, hsp = ss.hsp_queries(ss.SqL2(), _db(X), _db(np.zeros((1, 3), np.float32)), ids, d)
    np.testing.assert_array_equal(hi[:, 0], [2, 4, 1, 0])
    np.testing.assert_array_equal(hd[:, 0], np.array([1.0, 1.0, 4.0, np.inf], np.float32))
    assert len(hsp[0][0]) == 3


@pytest.mark.gpu
@pytest.mark.parametrize("metric,eps,n,blocksize,filterblocks", [
    ("SqL2", 0.01, 10000, 64, True), ("SqL2", 0.01, 9000, 1000, True),
    ("SqL2", 0.01, 9000, 700, False), ("L2", 0.1, 9000, 256, True),
    ("NormCosine", 0.004, 9000, 200, True)])
def test_neardup_several_batches(oracle, metric, eps, n, blocksize, filterblocks):
    """Long enough that blocks are searched in several batches (about 4096 objects each)
    against the net as of each batch start, plus the centers added since."""
    ss = _ss()
    X = _data(n, 4, 94, metric)
    X[n // 2:n // 2 + 300] = X[:300] + np.float32(1e-5)
    R = ss.neardup(_dist(ss, metric), _db(X), eps, blocksize=blocksize, filterblocks=filterblocks)
    M, L, D, centers, cost = oracle.neardup(metric, X, eps, blocksize=blocksize,
                                            filterblocks=filterblocks, nthreads=8)
    np.testing.assert_array_equal(R["map"], M)
    np.testing.assert_array_equal(R["nn"], L)
    np.testing.assert_array_equal(R["dist"], D)
    assert R["costdists"] == cost
    assert (R["nn"][n // 2:n // 2 + 300] != np.arange(n // 2 + 1, n // 2 + 301)).all()


@pytest.mark.gpu
@pytest.mark.parametrize("metric,eps,blocksize,pb", [("SqL2", 0.003, 200, 32), ("NormCosine", 2e-4, 256, 64)])
def test_neardup_recall_below_one_matches_oracle(oracle, metric, eps, blocksize, pb):
    """neardup with recall < 1 (neardup.jl:71-76): the net is a SearchGraph grown by
    append_items! of each block's new centers and searched with KnnSorted(8); the GPU
    (ssg_neardup_graph) equals the oracle's restatement in map, nn, dist and costdists, and
    the returned idx is the net's graph over the centers."""
    ss = _ss()
    X = _data(2400, 3, 131, metric)
    X[1500:1600] = X[:100] + np.float32(1e-5)
    R = ss.neardup(_dist(ss, metric), _db(X), eps, blocksize=blocksize, recall=0.9,
                   parallel_block=pb, seed=3)
    M, L, D, centers, cost = oracle.neardup(metric, X, eps, blocksize=blocksize, nthreads=8,
                                            graph=dict(parallel_block=pb, opt_minrecall=0.9, seed=3))
    assert len(centers) > 64
    np.testing.assert_array_equal(R["map"], M)
    np.testing.assert_array_equal(R["nn"], L)
    np.testing.assert_array_equal(R["dist"], D)
    assert R["costdists"] == cost
    # idx: index! over the first block's centers, then append_items! of each later block's
    # (the stage ends are the net sizes at the block boundaries)
    G = R["idx"]
    blk = (M.astype(np.int64) - 1) // blocksize
    ends = [int(e) for e in np.nonzero(np.diff(blk))[0] + 1] + [len(M)]
    Xc = X[M.astype(np.int64) - 1]
    g = oracle.build_graph(metric, Xc, parallel_block=pb, opt_minrecall=0.9, seed=3, stages=ends)
    np.testing.assert_array_equal(G.adj.offset, g["offs"][1:])
    np.testing.assert_array_equal(G.adj.end_point, g["adj"])
    np.testing.assert_array_equal(G.hints, g["hints"])
