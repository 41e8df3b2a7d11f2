"""The drop-in boundary beyond the default queue and distance: the queue-vector
searchbatch! with KnnSorted and KnnHeap queues and k up to n + 5
(test/testparallelexhaustive.jl:21-64 restated), exact exhaustive k > 32 (the
one-block-per-query scan), Cosine (src/dist/cos.jl:103), and graphs given as AdjList or
AdjDict (src/adj/*.jl).  Exhaustive ids and distances are bit-exact against the oracle;
heap queues are compared as id sets, as the reference's own test does."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ss():
    import similaritysearch as ss
    return ss


@pytest.mark.parametrize("qtype", ["KnnSorted", "KnnHeap"])
@pytest.mark.parametrize("k", [1, 4, 300, 305])
def test_parallel_exhaustive_queue_vector(oracle, qtype, k):
    """n = 300, dim 6, 40 queries; one queue of capacity min(k, n) per query."""
    ss = _ss()
    rng = np.random.default_rng(0)
    n = 300
    X = rng.random((n, 6), dtype=np.float32)
    Q = rng.random((40, 6), dtype=np.float32)
    T = getattr(ss, qtype)
    pex = ss.ParallelExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    knns = [ss.knnqueue(T, min(k, n)) for _ in range(len(Q))]
    ss.searchbatch_queues(pex, ss.GenericContext(), ss.MatrixDatabase.from_rows(Q), knns)
    gi, gd = oracle.exhaustive("SqL2", X, Q, min(k, n))
    for j, r in enumerate(knns):
        assert len(r) == min(k, n)
        ids, ds = r.sortitems()
        assert set(ids.tolist()) == set(gi[j].tolist())
        np.testing.assert_array_equal(np.sort(ds), gd[j])
        if qtype == "KnnSorted":
            np.testing.assert_array_equal(ids, gi[j])


@pytest.mark.parametrize("n,dim,k,grid", [
    (300, 6, 305, False),
    (5000, 16, 33, True),
    (5000, 16, 100, True),
    (4000, 128, 500, False),
    (6000, 32, 3840, True),
])
def test_exhaustive_large_k_bit_exact(oracle, n, dim, k, grid):
    """k beyond the 32 candidates stage 1 keeps per slice: first k in (d, id) order,
    ties to the lower id (integer grid rows tie often), 0 / +Inf past n."""
    ss = _ss()
    rng = np.random.default_rng(n + k)
    if grid:
        X = rng.integers(0, 3, (n, dim)).astype(np.float32)
        Q = rng.integers(0, 3, (37, dim)).astype(np.float32)
    else:
        X = rng.random((n, dim), dtype=np.float32)
        Q = rng.random((37, dim), dtype=np.float32)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    ids, d = ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(Q), k)
    gi, gd = oracle.exhaustive("SqL2", X, Q, k, nthreads=8)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)


def test_allknn_exhaustive_large_k(oracle):
    ss = _ss()
    X = np.random.default_rng(61).random((1500, 24), dtype=np.float32)
    E = ss.ExhaustiveSearch(ss.L2, ss.MatrixDatabase.from_rows(X))
    ids, d = ss.allknn(E, ss.GenericContext(), 40)
    gi, gd = oracle.exhaustive("L2", X, X, 40, nthreads=8)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)


def test_cosine_exhaustive_and_graph(oracle):
    """Cosine over rows of widely varying norms equals NormCosine over the oracle's
    normalised rows, for exhaustive search and for beam search on the same graph."""
    ss = _ss()
    rng = np.random.default_rng(62)
    X = (rng.standard_normal((3000, 48)) * rng.uniform(0.1, 50, (3000, 1))).astype(np.float32)
    Q = (rng.standard_normal((120, 48)) * rng.uniform(0.1, 50, (120, 1))).astype(np.float32)
    UX, UQ = oracle.unit_rows(X), oracle.unit_rows(Q)
    E = ss.ExhaustiveSearch(ss.Cosine, ss.MatrixDatabase.from_rows(X))
    ids, d = ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(Q), 12)
    gi, gd = oracle.exhaustive("NormCosine", UX, UQ, 12)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    g = oracle.build_graph("NormCosine", UX, parallel_block=64, seed=3)
    G = ss.SearchGraph(ss.Cosine, ss.MatrixDatabase.from_rows(X),
                       ss.StaticAdjList.from_csr(g["offs"], g["adj"]), g["hints"],
                       ss.BeamSearch(12, 1.05, 10**6))
    ctx = ss.SearchGraphContext()
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
    bi, bd, bcd, bcb = oracle.beamsearch("NormCosine", UX, g["offs"], g["adj"], g["hints"], UQ, 10,
                                         12, 1.05, 10**6)
    np.testing.assert_array_equal(ids.T, bi)
    np.testing.assert_array_equal(d.T, bd)
    np.testing.assert_array_equal(ctx.costdists, bcd)


@pytest.mark.parametrize("form", ["AdjList", "AdjDict"])
def test_graph_from_adjacency_lists(oracle, form):
    """A SearchGraph given its lists as AdjList (the reference's default, SearchGraph.jl:164)
    or AdjDict searches exactly like the same lists in StaticAdjList form; queue-vector
    searches fill KnnHeap and KnnSorted queues with the same results."""
    ss = _ss()
    rng = np.random.default_rng(63)
    X = rng.random((2500, 8), dtype=np.float32)
    Q = rng.random((90, 8), dtype=np.float32)
    g = oracle.build_graph("SqL2", X, parallel_block=64, seed=4)
    offs, adj = g["offs"], g["adj"]
    lists = [adj[offs[i]:offs[i + 1]].tolist() for i in range(len(X))]
    A = ss.AdjList(lists) if form == "AdjList" else \
        ss.AdjDict({i + 1: l for i, l in enumerate(lists) if l}, len(X))
    G = ss.SearchGraph(ss.SqL2, ss.MatrixDatabase.from_rows(X), A, g["hints"],
                       ss.BeamSearch(g["bsize"], g["delta"], g["maxvisits"]))
    ctx = ss.SearchGraphContext()
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
    bi, bd, bcd, bcb = oracle.beamsearch("SqL2", X, offs, adj, g["hints"], Q, 10, g["bsize"],
                                         g["delta"], g["maxvisits"])
    np.testing.assert_array_equal(ids.T, bi)
    np.testing.assert_array_equal(d.T, bd)
    np.testing.assert_array_equal(ctx.costdists, bcd)
    knns = [ss.knnqueue(ss.KnnHeap if j % 2 else ss.KnnSorted, 1 + j % 10) for j in range(len(Q))]
    ss.searchbatch_queues(G, ss.SearchGraphContext(), ss.MatrixDatabase.from_rows(Q), knns)
    ks = np.array([1 + j % 10 for j in range(len(Q))], np.int32)
    oi, od, _, _ = oracle.beamsearch("SqL2", X, offs, adj, g["hints"], Q, 10, g["bsize"], g["delta"],
                                     g["maxvisits"], kq=ks)
    for j, r in enumerate(knns):
        ri, rd = r.sortitems()
        assert set(ri.tolist()) == set(oi[j, :ks[j]].tolist())
        np.testing.assert_array_equal(np.sort(rd), od[j, :ks[j]])


@pytest.mark.parametrize("dim", [136, 150, 176, 200, 250, 300, 512, 700, 1000, 1024])
def test_any_dim_up_to_1024(oracle, dim):
    """Dimensions that pad to 192, 256, 384, 512, 768 or 1024 (rows padded with zeros; the
    canonical distance's 16 runs cover the padded row, in the GPU and the oracle alike):
    exhaustive k-NN (above 256 dims the one-block-per-query scan), beam search over an
    oracle-built graph, and rerank!, bit-exact."""
    ss = _ss()
    rng = np.random.default_rng(dim)
    X = rng.standard_normal((1500, dim)).astype(np.float32)
    Q = rng.standard_normal((60, dim)).astype(np.float32)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    ids, d = ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(Q), 10)
    gi, gd = oracle.exhaustive("SqL2", X, Q, 10)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    g = oracle.build_graph("SqL2", X, parallel_block=64, seed=2)
    G = ss.SearchGraph(ss.SqL2, ss.MatrixDatabase.from_rows(X),
                       ss.StaticAdjList.from_csr(g["offs"], g["adj"]), g["hints"],
                       ss.BeamSearch(g["bsize"], g["delta"], g["maxvisits"]))
    ctx = ss.SearchGraphContext()
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
    bi, bd, bcd, bcb = oracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 10,
                                         g["bsize"], g["delta"], g["maxvisits"])
    np.testing.assert_array_equal(ids.T, bi)
    np.testing.assert_array_equal(d.T, bd)
    np.testing.assert_array_equal(ctx.costdists, bcd)
    I = np.asfortranarray(rng.integers(1, 1501, (12, 60)).astype(np.uint32))
    D = np.zeros(I.shape, np.float32, order="F")
    ri, rd = oracle.rerank("SqL2", X, Q, np.ascontiguousarray(I.T), np.ascontiguousarray(D.T))
    ss.rerank_(ss.SqL2, ss.MatrixDatabase.from_rows(X), ss.MatrixDatabase.from_rows(Q), I, D)
    np.testing.assert_array_equal(I.T, ri)
    np.testing.assert_array_equal(D.T, rd)


def test_bf16_form_above_256_dims_is_refused(oracle):
    ss = _ss()
    rng = np.random.default_rng(5)
    X = rng.standard_normal((300, 300)).astype(np.float32)
    g = oracle.build_graph("SqL2", X, parallel_block=64, seed=2)
    G = ss.SearchGraph(ss.SqL2, ss.MatrixDatabase.from_rows(X),
                       ss.StaticAdjList.from_csr(g["offs"], g["adj"]), g["hints"],
                       ss.BeamSearch(g["bsize"], g["delta"], g["maxvisits"]))
    with pytest.raises(Exception):
        ss.searchbatch(G, ss.SearchGraphContext(precision="bf16", rerank=14),
                       ss.MatrixDatabase.from_rows(X[:4]), 10)


def test_dim_above_1024_is_refused():
    ss = _ss()
    X = np.zeros((100, 1025), np.float32)
    with pytest.raises(Exception):
        E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
        ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(X[:2]), 1)
