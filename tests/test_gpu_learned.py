"""GPU checks of the landmark and precision forms against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ss():
    import similaritysearch as ss
    return ss


@pytest.mark.parametrize("metric,n,dim,k,start", [
    ("SqL2", 3000, 8, 64, 1),
    ("SqL2", 5000, 32, 200, 17),
    ("L2", 2000, 128, 50, 5),
    ("NormCosine", 1500, 96, 40, 3),
])
def test_fft_matches_oracle(oracle, metric, n, dim, k, start):
    ss = _ss()
    rng = np.random.default_rng(n + k)
    X = rng.random((n, dim), dtype=np.float32)
    if metric == "NormCosine":
        X /= np.linalg.norm(X, axis=1, keepdims=True)
    dist = {"SqL2": ss.SqL2, "L2": ss.L2, "NormCosine": ss.NormCosine}[metric]
    c, nn, nnd, eps = ss.fft(dist, ss.MatrixDatabase.from_rows(X), k, start=start)
    oc, onn, onnd, oeps = oracle.fft(metric, X, k, start=start)
    np.testing.assert_array_equal(c, oc)
    np.testing.assert_array_equal(nn, onn)
    np.testing.assert_array_equal(nnd, onnd)
    np.testing.assert_array_equal(eps, oeps)


def test_bf16_form_returns_exact_distances_and_high_recall(oracle):
    ss = _ss()
    rng = np.random.default_rng(71)
    X = rng.random((4000, 32), dtype=np.float32)
    Q = rng.random((300, 32), dtype=np.float32)
    g = oracle.build_graph("SqL2", X, parallel_block=64, seed=9)
    adj = ss.StaticAdjList.from_csr(g["offs"], g["adj"])
    G = ss.SearchGraph(ss.SqL2, ss.MatrixDatabase.from_rows(X), adj, g["hints"],
                       ss.BeamSearch(16, 1.05, 10**6))
    ctx = ss.SearchGraphContext()
    ids32, _ = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
    ids16, d16 = ss.searchbatch(G, ss.SearchGraphContext(precision="bf16", rerank=20),
                                ss.MatrixDatabase.from_rows(Q), 10)
    gi, _ = oracle.exhaustive("SqL2", X, Q, 10)
    r32 = oracle.macrorecall(gi, ids32.T)
    r16 = oracle.macrorecall(gi, ids16.T)
    assert r16 >= r32 - 0.01
    for j in range(len(Q)):
        ex = [oracle.dist("SqL2", X[i - 1], Q[j]) for i in ids16[:, j]]
        np.testing.assert_array_equal(d16[:, j], np.array(ex, np.float32))
        assert np.all(np.diff(d16[:, j]) >= 0)


def test_mlp_gradients_match_oracle(oracle):
    import mlp_oracle as M
    from similaritysearch.learned import Mlp
    rng = np.random.default_rng(3)
    D, H1, H2, L, NB, B = 16, 48, 40, 70, 5, 96
    mlp = Mlp(D, H1, H2, L, NB, seed=4)
    W, b = [], []
    for l in range(3):
        w, bb = mlp.get(l)
        bb = (rng.standard_normal(bb.shape) * 0.1).astype(np.float32)
        mlp.set(l, w, bb)
        W.append(w.astype(np.float64))
        b.append(bb.astype(np.float64))
    X = rng.standard_normal((B, D)).astype(np.float32)
    yl = rng.integers(0, L, B).astype(np.int32)
    yb = rng.integers(0, NB, B).astype(np.int32)
    loss = mlp.step(X, yl, yb, bucket_weight=0.5, update=False)
    rl, rgW, rgb = M.forward_backward(W, b, X, yl, yb, L, NB, 0.5, half=True)
    assert loss == pytest.approx(rl, rel=2e-3)
    for l in range(3):
        gW, gb = mlp.grads(l)
        scale = np.abs(rgW[l]).max()
        assert np.abs(gW - rgW[l]).max() <= 2e-2 * scale
        assert np.abs(gb - rgb[l]).max() <= 2e-2 * np.abs(rgb[l]).max() + 1e-7


def test_mlp_training_learns_clusters():
    """The trained network beats chance by a wide margin on a separable task."""
    from similaritysearch.learned import Mlp
    rng = np.random.default_rng(5)
    D, L, NB, N = 32, 50, 4, 6000
    C0 = rng.standard_normal((L, D)).astype(np.float32) * 3
    yl = rng.integers(0, L, N).astype(np.int32)
    X = (C0[yl] + rng.standard_normal((N, D)).astype(np.float32)).astype(np.float32)
    yb = (yl % NB).astype(np.int32)
    mlp = Mlp(D, 64, 64, L, NB, seed=6)
    first = mlp.step(X[:512], yl[:512], yb[:512], lr=3e-3)
    for ep in range(8):
        for s in range(0, N - 512 + 1, 512):
            last = mlp.step(X[s:s + 512], yl[s:s + 512], yb[s:s + 512], lr=3e-3)
    assert last < 0.5 * first
    # accuracy through the FP32 weights in numpy (independent forward)
    Ws = [mlp.get(l) for l in range(3)]
    h = np.maximum(X @ Ws[0][0].T + Ws[0][1], 0)
    h = np.maximum(h @ Ws[1][0].T + Ws[1][1], 0)
    z = h @ Ws[2][0].T + Ws[2][1]
    assert (z[:, :L].argmax(1) == yl).mean() > 0.9


def test_learned_search_recall_and_fallback(oracle):
    import similaritysearch as ss
    from similaritysearch.learned import LearnedSearch
    from similaritysearch import datasets
    X = datasets.gmm(20000, 32, seed=31)
    Qt = datasets.gmm(8000, 32, seed=32)
    Q = datasets.gmm(500, 32, seed=33)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=128, seed=3)
    ctx = ss.SearchGraphContext()
    LS = LearnedSearch(G, nlandmarks=256, hidden=64, buckets=(2, 4, 8, 16, 32, 64, 128, 256))
    losses = LS.fit(Qt, ctx, epochs=6, batch=256)
    assert losses[-1] < losses[0]
    gi, _ = oracle.exhaustive("SqL2", X, Q, 10)
    ids, d = LS.search(ctx, ss.MatrixDatabase.from_rows(Q), 10, nhints=2, shift=1)
    rec = oracle.macrorecall(gi, ids.T)
    assert rec > 0.8
    # every returned distance is the exact canonical distance
    for j in range(0, 500, 50):
        ex = [oracle.dist("SqL2", X[i - 1], Q[j]) for i in ids[:, j]]
        np.testing.assert_array_equal(d[:, j], np.array(ex, np.float32))
    # min_conf above every confidence: all queries take the plain path (global hints) in
    # the BF16 form; same ids as the BF16 plain search with the same beam size
    ids_fb, _ = LS.search(ctx, ss.MatrixDatabase.from_rows(Q), 10, min_conf=2.0, fallback_bsize=8)
    G.algo = ss.BeamSearch(8, G.algo.delta, 10**6)
    ids_plain, _ = ss.searchbatch(G, ss.SearchGraphContext(precision="bf16", rerank=20),
                                  ss.MatrixDatabase.from_rows(Q), 10)
    np.testing.assert_array_equal(ids_fb, ids_plain)


def test_learned_network_gain_over_plain_search(oracle):
    """What the trained network buys: at the recall the learned form reaches, the plain
    search of the same form (BF16 gathers, exact rerank of 20) needs far more distance
    evaluations for the same recall (1,000 queries, 20,000 x 32 mixture)."""
    import similaritysearch as ss
    from similaritysearch.learned import LearnedSearch
    from similaritysearch import datasets
    X = datasets.gmm(20000, 32, seed=31)
    Qt = datasets.gmm(8000, 32, seed=32)
    Q = datasets.gmm(1000, 32, seed=33)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=128, seed=3)
    ctx = ss.SearchGraphContext()
    LS = LearnedSearch(G, nlandmarks=256, hidden=64)
    LS.fit(Qt, ctx, epochs=8, batch=256)
    gi, _ = oracle.exhaustive("SqL2", X, Q, 10)
    Qdb = ss.MatrixDatabase.from_rows(Q)
    ids, _ = LS.search(ctx, Qdb, 10, nhints=1, rerank=20, quantile=0.7, delta=1.1)
    l_rec, l_cost = oracle.macrorecall(gi, ids.T), float(ctx.costdists.mean())
    assert l_rec >= 0.95
    plain_cost = None
    for b in (2, 4, 6, 8, 12, 16, 24, 32, 48, 64):
        G.algo = ss.BeamSearch(b, 1.1, 10**6)
        c2 = ss.SearchGraphContext(precision="bf16", rerank=20, beamcap=64)
        pids, _ = ss.searchbatch(G, c2, Qdb, 10)
        if oracle.macrorecall(gi, pids.T) >= l_rec - 0.005:
            plain_cost = float(c2.costdists.mean())
            break
    assert plain_cost is not None
    assert l_cost < 0.8 * plain_cost, (l_rec, l_cost, plain_cost)
