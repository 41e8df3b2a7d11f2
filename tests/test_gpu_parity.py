"""GPU parity: libssg_hip.so (through its C ABI) against the CPU oracle on the same
seeded inputs.  Integer results (ids, cost counters) must match bit-exactly; the
distances are the canonical distance, which both sides evaluate with the same
expression tree, so they must match exactly as well (the tolerance the contract
allows is 1e-5 relative; equality is asserted because it holds).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ss():
    import similaritysearch as ss
    return ss


def _uniform(n, dim, seed):
    return np.random.default_rng(seed).random((n, dim), dtype=np.float32)


def _unit(n, dim, seed):
    x = np.random.default_rng(seed).standard_normal((n, dim)).astype(np.float32)
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


def _exh(ss, oracle, metric, X, Q, k):
    dist = {"SqL2": ss.SqL2, "L2": ss.L2, "NormCosine": ss.NormCosine}[metric]
    E = ss.ExhaustiveSearch(dist, ss.MatrixDatabase.from_rows(X))
    ids, d = ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(Q), k)
    gi, gd = oracle.exhaustive(metric, X, Q, k)
    return ids.T, d.T, gi, gd


@pytest.mark.parametrize("metric,dim,n,m,k", [
    ("SqL2", 32, 3000, 100, 10),
    ("SqL2", 4, 777, 65, 1),
    ("L2", 128, 2500, 70, 32),
    ("NormCosine", 96, 2000, 64, 10),
    ("SqL2", 100, 1500, 33, 7),
])
def test_exhaustive_matches_oracle(oracle, metric, dim, n, m, k):
    ss = _ss()
    X = _unit(n, dim, 11) if metric == "NormCosine" else _uniform(n, dim, 11)
    Q = _unit(m, dim, 12) if metric == "NormCosine" else _uniform(m, dim, 12)
    ids, d, gi, gd = _exh(ss, oracle, metric, X, Q, k)
    np.testing.assert_array_equal(ids, gi)
    np.testing.assert_array_equal(d, gd)


@pytest.mark.parametrize("metric,dim,n,m,k,order", [
    ("SqL2", 32, 40000, 300, 10, "random"),
    ("SqL2", 32, 40000, 300, 1, "far_first"),
    ("L2", 128, 36000, 200, 32, "random"),
    ("NormCosine", 96, 50000, 150, 10, "far_first"),
    ("SqL2", 16, 40000, 257, 10, "duplicates"),
])
def test_exhaustive_seeded_thresholds(oracle, metric, dim, n, m, k, order):
    """Short slices (n >= 32768 over many slices) take a seed pass over the first n/16 rows
    whose k-th best score starts each query's threshold.  Bit-exact whatever the rows'
    order: far_first puts the rows farthest from every query first (a weak seed), and
    duplicates repeats rows so that distances tie at the threshold."""
    ss = _ss()
    X = _unit(n, dim, 13) if metric == "NormCosine" else _uniform(n, dim, 13)
    Q = _unit(m, dim, 14) if metric == "NormCosine" else _uniform(m, dim, 14)
    if order == "far_first":
        c = Q.mean(axis=0)
        X = X[np.argsort(-((X - c) ** 2).sum(axis=1), kind="stable")]
    elif order == "duplicates":
        X = np.repeat(X[: n // 8], 8, axis=0)
        Q = X[np.random.default_rng(15).integers(0, n, m)] + np.float32(1e-3)
    X = np.ascontiguousarray(X, np.float32)
    Q = np.ascontiguousarray(Q, np.float32)
    ids, d, gi, gd = _exh(ss, oracle, metric, X, Q, k)
    np.testing.assert_array_equal(ids, gi)
    np.testing.assert_array_equal(d, gd)


def test_exhaustive_fewer_objects_than_k(oracle):
    ss = _ss()
    X = _uniform(5, 16, 3)
    Q = _uniform(9, 16, 4)
    ids, d, gi, gd = _exh(ss, oracle, "SqL2", X, Q, 8)
    np.testing.assert_array_equal(ids, gi)
    np.testing.assert_array_equal(d, gd)
    assert (ids[:, 5:] == 0).all() and np.isinf(d[:, 5:]).all()


def test_exhaustive_ties_take_lower_id(oracle):
    """Integer grid data: many equal distances, the certificate fails and the exact
    rescan must reproduce the ascending-scan tie rule."""
    ss = _ss()
    rng = np.random.default_rng(5)
    X = rng.integers(0, 3, (4000, 8)).astype(np.float32)
    Q = rng.integers(0, 3, (40, 8)).astype(np.float32)
    ids, d, gi, gd = _exh(ss, oracle, "SqL2", X, Q, 10)
    np.testing.assert_array_equal(ids, gi)
    np.testing.assert_array_equal(d, gd)


def test_allknn_exhaustive_self_first(oracle):
    ss = _ss()
    X = _uniform(1200, 32, 8)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    ids, d = ss.allknn(E, ss.GenericContext(), 5)
    gi, gd = oracle.exhaustive("SqL2", X, X, 5)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    assert (ids[0, :] == np.arange(1, 1201)).all()


@pytest.fixture(scope="module")
def small_graph(oracle):
    X = _uniform(3000, 8, 21)
    g = oracle.build_graph("SqL2", X, parallel_block=64, seed=7)
    return X, g


def _graph_index(ss, X, g, metric=None, algo=None):
    adj = ss.StaticAdjList.from_csr(g["offs"], g["adj"])
    return ss.SearchGraph(metric or ss.SqL2, ss.MatrixDatabase.from_rows(X), adj, g["hints"],
                          algo or ss.BeamSearch(g["bsize"], g["delta"], g["maxvisits"]))


@pytest.mark.parametrize("bsize,delta,maxvisits,beamcap,k", [
    (4, 1.0, 10**6, 32, 10),
    (16, 1.1, 10**6, 32, 10),
    (2, 0.9, 10**6, 32, 1),
    (100, 1.05, 10**6, 256, 50),
    (8, 1.0, 40, 32, 10),
    (40, 1.0, 10**6, 32, 64),
])
def test_beamsearch_matches_oracle(oracle, small_graph, bsize, delta, maxvisits, beamcap, k):
    ss = _ss()
    X, g = small_graph
    Q = _uniform(200, 8, 22)
    algo = ss.BeamSearch(bsize, delta, maxvisits)
    G = _graph_index(ss, X, g, algo=algo)
    ctx = ss.SearchGraphContext(beamcap=beamcap)
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), k)
    gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, k, bsize,
                                         delta, maxvisits, beamcap=beamcap)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    np.testing.assert_array_equal(ctx.costdists, gcd)
    np.testing.assert_array_equal(ctx.costblocks, gcb)


def test_beamsearch_per_query_k_bsize_hints(oracle, small_graph):
    ss = _ss()
    X, g = small_graph
    m = 150
    Q = _uniform(m, 8, 23)
    rng = np.random.default_rng(24)
    ks = rng.choice([1, 10, 50], m).astype(np.int32)
    bs = rng.choice([16, 32, 64, 128, 256], m).astype(np.int32)
    hints = [rng.integers(1, len(X) + 1, rng.integers(0, 6)).astype(np.uint32) for _ in range(m)]
    G = _graph_index(ss, X, g)
    ctx = ss.SearchGraphContext(beamcap=256)
    out = ss.searchbatch_knns(G, ctx, ss.MatrixDatabase.from_rows(Q), ks, bsizes=bs, hints=hints)
    stride = max(1, max(len(h) for h in hints))
    H = np.zeros((m, stride), np.uint32)
    cnt = np.zeros(m, np.int32)
    for j, h in enumerate(hints):
        H[j, :len(h)] = h
        cnt[j] = len(h)
    gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 1,
                                         g["bsize"], g["delta"], g["maxvisits"], beamcap=256,
                                         kq=ks, bq=bs, qhints=H, qh_count=cnt)
    for j in range(m):
        np.testing.assert_array_equal(out[j][0], gi[j, :ks[j]])
        np.testing.assert_array_equal(out[j][1], gd[j, :ks[j]])
    np.testing.assert_array_equal(ctx.costdists, gcd)
    np.testing.assert_array_equal(ctx.costblocks, gcb)


def test_beamsearch_tiny_graph_brute_force(oracle):
    """n < 64 takes the brute-force branch (beamsearch.jl:94-101)."""
    ss = _ss()
    X = _uniform(50, 8, 31)
    g = oracle.build_graph("SqL2", X, parallel_block=1, seed=3)
    Q = _uniform(20, 8, 32)
    G = _graph_index(ss, X, g)
    ctx = ss.SearchGraphContext()
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 6)
    gi, gd = oracle.exhaustive("SqL2", X, Q, 6)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    assert (ctx.costdists == 50).all()


def test_allknn_graph_matches_oracle(oracle, small_graph):
    ss = _ss()
    X, g = small_graph
    G = _graph_index(ss, X, g)
    ids, d = ss.allknn(G, ss.SearchGraphContext(), 8)
    gi, gd = oracle.allknn_graph("SqL2", X, g["offs"], g["adj"], 8, g["bsize"], g["delta"],
                                 g["maxvisits"])
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)


def test_rerank_matches_oracle(oracle):
    ss = _ss()
    X = _uniform(900, 32, 41)
    Q = _uniform(50, 32, 42)
    rng = np.random.default_rng(43)
    ids = rng.integers(1, 901, (50, 16)).astype(np.uint32)
    for j in range(50):
        ids[j, rng.integers(5, 17):] = 0
    d = np.zeros((50, 16), np.float32)
    gi, gd = oracle.rerank("SqL2", X, Q, ids, d)
    I = np.asfortranarray(ids.T)
    D = np.asfortranarray(d.T)
    ss.rerank_(ss.SqL2, ss.MatrixDatabase.from_rows(X), ss.MatrixDatabase.from_rows(Q), I, D)
    np.testing.assert_array_equal(I.T, gi)
    np.testing.assert_array_equal(D.T, gd)


def test_errors_are_argument_errors():
    ss = _ss()
    X = _uniform(100, 8, 51)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    with pytest.raises(ValueError):
        ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(np.zeros((0, 8), np.float32)), 3)
    with pytest.raises(ValueError):
        ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(_uniform(4, 7, 1)), 3)


@pytest.mark.parametrize("n,dim,pb,seed", [(3000, 8, 64, 7), (1500, 16, 200, 2)])
def test_gpu_index_build_matches_oracle(oracle, n, dim, pb, seed):
    """index! on the GPU produces exactly the oracle's graph, hints and tuned BeamSearch."""
    ss = _ss()
    X = _uniform(n, dim, 100 + seed)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=pb, seed=seed)
    g = oracle.build_graph("SqL2", X, parallel_block=pb, seed=seed)
    np.testing.assert_array_equal(G.adj.offset, g["offs"][1:])
    np.testing.assert_array_equal(G.adj.end_point, g["adj"])
    np.testing.assert_array_equal(G.hints, g["hints"])
    assert (G.algo.bsize, np.float32(G.algo.delta), G.algo.maxvisits) == \
        (g["bsize"], np.float32(g["delta"]), g["maxvisits"])


def test_gpu_optimize_matches_oracle(oracle, small_graph):
    ss = _ss()
    X, g = small_graph
    Q = _uniform(64, 8, 61)
    G = _graph_index(ss, X, g)
    algo, rec = G.optimize(ss.SearchGraphContext(), 0.9, ss.MatrixDatabase.from_rows(Q))
    o = oracle.optimize_minrecall("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 10, 0.9,
                                  cur_maxvisits=g["maxvisits"])
    assert (algo.bsize, np.float32(algo.delta), algo.maxvisits) == \
        (o["bsize"], np.float32(o["delta"]), o["maxvisits"])
    assert rec == pytest.approx(o["recall"])


@pytest.mark.parametrize("stages", [[1000, 1700, 3000], [640, 1280, 3000], [300, 301, 2999, 3000]])
def test_gpu_append_items_graph_matches_oracle(oracle, stages):
    """append_items! on a SearchGraph (insertions.jl:27-34): index! continues from
    length(index) + 1 over the new objects; the graph, hints and tuned BeamSearch after each
    append equal the oracle's build with the same stages; searches over the grown graph
    match the oracle's."""
    ss = _ss()
    X = _uniform(3000, 8, 107)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X[:stages[0]]), parallel_block=64,
                             seed=7)
    for a, b in zip(stages[:-1], stages[1:]):
        G.append_items(ss.SearchGraphContext(), ss.MatrixDatabase.from_rows(X[a:b]))
    g = oracle.build_graph("SqL2", X, parallel_block=64, seed=7, stages=stages)
    np.testing.assert_array_equal(G.adj.offset, g["offs"][1:])
    np.testing.assert_array_equal(G.adj.end_point, g["adj"])
    np.testing.assert_array_equal(G.hints, g["hints"])
    assert (G.algo.bsize, np.float32(G.algo.delta), G.algo.maxvisits) == \
        (g["bsize"], np.float32(g["delta"]), g["maxvisits"])
    Q = _uniform(200, 8, 108)
    ctx = ss.SearchGraphContext()
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
    bi, bd, bcd, bcb = oracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 10,
                                         g["bsize"], g["delta"], g["maxvisits"])
    np.testing.assert_array_equal(ids.T, bi)
    np.testing.assert_array_equal(d.T, bd)
    np.testing.assert_array_equal(ctx.costdists, bcd)


def test_gpu_append_items_needs_an_index_built_here(oracle, small_graph):
    """A graph uploaded as CSR (not built by index! on the handle) cannot be continued."""
    ss = _ss()
    X, g = small_graph
    G = _graph_index(ss, X, g)
    with pytest.raises(Exception):
        G.append_items(ss.SearchGraphContext(), ss.MatrixDatabase.from_rows(X[:10]))
