"""Pins of the CPU oracle (oracle/ssg_oracle.c) against the reference's known answers,
closed forms, invariants and brute force on small inputs.  None of these checks
re-types the oracle's formulas: each compares against an independent formulation
(FP64 numpy, sort-based golds, hand-traced searches, structural invariants)."""
import numpy as np
import pytest


# ---------------------------------------------------------------- distances
def test_distance_known_answers(oracle):
    # test/testdistances.jl:34-76
    u = np.array([1, 2, 3], np.float32)
    v = np.array([4, 6, 8], np.float32)
    assert oracle.dist("SqL2", u, v) == 50.0
    assert oracle.dist("L2", u, v) == np.float32(np.sqrt(50.0))
    assert oracle.dist("L2", u, u) == 0.0
    p = np.array([3, 4], np.float32)
    q = np.zeros(2, np.float32)
    assert oracle.dist("L2", p, q) == 5.0
    assert oracle.dist("SqL2", p, q) == 25.0
    e = np.array([1, 0, 0], np.float32)
    f = np.array([0, 1, 0], np.float32)
    assert oracle.dist("SqL2", e, f) == 2.0
    assert abs(oracle.dist("L2", e, f) - np.sqrt(2.0)) < 1e-6
    assert oracle.dist("NormCosine", e, f) == 1.0
    assert oracle.dist("NormCosine", e, e) == 0.0


def test_padded_dim_rule(oracle):
    """The canonical distance's runs cover roundup(dim, 16) up to 128, else 192, 256, 384,
    512, 768 or 1024: the row sizes the GPU kernels are built for."""
    f = oracle.lib().or_padded_dim
    got = {d: f(d) for d in (1, 16, 17, 100, 128, 129, 150, 176, 192, 193, 240, 256, 257, 384,
                             385, 512, 513, 768, 769, 1024)}
    assert got == {1: 16, 16: 16, 17: 32, 100: 112, 128: 128, 129: 192, 150: 192, 176: 192,
                   192: 192, 193: 256, 240: 256, 256: 256, 257: 384, 384: 384, 385: 512,
                   512: 512, 513: 768, 768: 768, 769: 1024, 1024: 1024}


@pytest.mark.parametrize("dim", [1, 4, 15, 16, 17, 32, 96, 100, 128, 150, 200, 256, 300, 1000])
@pytest.mark.parametrize("metric", ["SqL2", "L2", "NormCosine"])
def test_distance_is_rounded_fp64(oracle, dim, metric):
    """canonical distance = float32 of the exact value (within one float32 ulp)."""
    rng = np.random.default_rng(dim)
    A = rng.standard_normal((200, dim)).astype(np.float32)
    B = rng.standard_normal((200, dim)).astype(np.float32)
    if metric == "NormCosine":
        A /= np.linalg.norm(A, axis=1, keepdims=True)
        B /= np.linalg.norm(B, axis=1, keepdims=True)
    a64, b64 = A.astype(np.float64), B.astype(np.float64)
    if metric == "NormCosine":
        exact = 1.0 - np.sum(a64 * b64, axis=1)
    else:
        exact = np.sum((a64 - b64) ** 2, axis=1)
        if metric == "L2":
            exact = np.sqrt(exact)
    got = np.array([oracle.dist(metric, A[i], B[i]) for i in range(200)], np.float64)
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - exact) <= ulp + 1e-12)


def test_dist_f64_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    a = rng.random(96, dtype=np.float32)
    b = rng.random(96, dtype=np.float32)
    assert oracle.dist_f64("SqL2", a, b) == pytest.approx(
        float(np.sum((a.astype(np.float64) - b) ** 2)), rel=1e-14)


# ---------------------------------------------------------------- KnnSorted
@pytest.mark.parametrize("k", [1, 7, 8, 12, 15, 67])
def test_knnsorted_against_sort_gold(oracle, k):
    """test/testresults.jl:57-85: after every push the queue equals the k first items of a
    stable sort of everything pushed so far."""
    rng = np.random.default_rng(k)
    n = 1000
    d = rng.random(n, dtype=np.float32)
    d[::7] = d[3]  # plenty of ties
    ids = np.arange(1, n + 1, dtype=np.uint32)
    for stop in (5, 50, 1000):
        got_i, got_d, acc = oracle.knn_run(k, np.zeros(stop, np.int8), ids[:stop], d[:stop])
        order = np.argsort(d[:stop], kind="stable")[:k]
        np.testing.assert_array_equal(got_i, ids[:stop][order])
        np.testing.assert_array_equal(got_d, d[:stop][order])


def test_knnsorted_full_queue_rejects_equal_maximum(oracle):
    ids = np.array([1, 2, 3, 4], np.uint32)
    d = np.array([0.1, 0.5, 0.5, 0.3], np.float32)
    gi, gd, acc = oracle.knn_run(2, np.zeros(4, np.int8), ids, d)
    np.testing.assert_array_equal(acc, [True, True, False, True])
    np.testing.assert_array_equal(gi, [1, 4])


def test_knnsorted_pop_min_and_refill(oracle):
    # pops shrink from the front; later pushes still land in order (compaction path)
    ops = np.array([0, 0, 0, 1, 0, 1, 0, 0], np.int8)
    ids = np.array([1, 2, 3, 0, 4, 0, 5, 6], np.uint32)
    d = np.array([0.3, 0.1, 0.2, 0, 0.15, 0, 0.05, 0.25], np.float32)
    gi, gd, acc = oracle.knn_run(3, ops, ids, d)
    # after pushes 1,2,3: [2 .1, 3 .2, 1 .3]; pop -> [3, 1]; push 4 (.15) -> [4, 3, 1];
    # pop -> [3, 1]; push 5 (.05) -> [5, 3, 1]; push 6 (.25): full, .25 < .3 -> [5, 3, 6]
    np.testing.assert_array_equal(gi, [5, 3, 6])
    np.testing.assert_array_equal(gd, np.array([0.05, 0.2, 0.25], np.float32))
    with pytest.raises(ValueError):
        oracle.knn_run(2, np.array([1], np.int8), np.zeros(1, np.uint32), np.zeros(1, np.float32))


# ---------------------------------------------------------------- exhaustive
def _exact_order(X, Q):
    d = ((Q.astype(np.float64)[:, None, :] - X.astype(np.float64)[None, :, :]) ** 2).sum(-1)
    return d, np.argsort(d, axis=1, kind="stable")


def test_exhaustive_against_fp64_brute_force(oracle):
    rng = np.random.default_rng(1)
    X = rng.random((3000, 32), dtype=np.float32)
    Q = rng.random((40, 32), dtype=np.float32)
    ids, dists = oracle.exhaustive("SqL2", X, Q, 10)
    d64, order = _exact_order(X, Q)
    for j in range(len(Q)):
        srt = np.sort(d64[j])
        near_tie = np.any(np.diff(srt[:11]) < 4 * np.spacing(np.float32(srt[10])))
        if not near_tie:
            np.testing.assert_array_equal(ids[j], order[j, :10] + 1)
        np.testing.assert_allclose(dists[j], d64[j, ids[j] - 1], rtol=1e-6)


def test_exhaustive_self_query_and_empty_slots(oracle):
    rng = np.random.default_rng(2)
    X = rng.random((50, 8), dtype=np.float32)
    ids, dists = oracle.exhaustive("SqL2", X, X[:10], 3)
    np.testing.assert_array_equal(ids[:, 0], np.arange(1, 11))
    assert (dists[:, 0] == 0).all()
    ids, dists = oracle.exhaustive("SqL2", X[:2], X[:3], 4)
    assert (ids[:, 2:] == 0).all() and np.isinf(dists[:, 2:]).all()


@pytest.mark.parametrize("nb", [1, 2, 3, 7, 64])
@pytest.mark.parametrize("k", [1, 4, 20])
def test_parallel_equals_sequential(oracle, nb, k):
    """test/testparallelexhaustive.jl:21-64 (exact equality: both use one distance)."""
    rng = np.random.default_rng(nb * 100 + k)
    X = rng.integers(0, 4, (500, 6)).astype(np.float32)  # many ties
    Q = rng.integers(0, 4, (30, 6)).astype(np.float32)
    a = oracle.exhaustive("SqL2", X, Q, k)
    b = oracle.parallel_exhaustive("SqL2", X, Q, k, nb)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_ties_resolve_to_lower_id(oracle):
    X = np.zeros((10, 4), np.float32)
    Q = np.ones((1, 4), np.float32)
    ids, _ = oracle.exhaustive("SqL2", X, Q, 3)
    np.testing.assert_array_equal(ids[0], [1, 2, 3])


# ---------------------------------------------------------------- beam search
def _chain(n):
    """points 1..n on a line, node i linked to i-1 and i+1 (1-based)."""
    X = np.arange(1, n + 1, dtype=np.float32)[:, None] * np.ones((1, 4), np.float32) / 2.0
    lists = []
    for i in range(1, n + 1):
        lists.append([j for j in (i - 1, i + 1) if 1 <= j <= n])
    offs = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    adj = np.array([x for l in lists for x in l], np.uint32)
    return X, offs, adj


def test_beam_walk_along_a_chain(oracle):
    """Hand trace: bsize 1, Δ 1, k 1, hint node 1, query at node 50 + 0.3: the walk
    expands nodes 1..50, evaluates nodes 2..51 once each, and returns node 50."""
    X, offs, adj = _chain(70)
    q = (X[49] + 0.15)[None, :]
    ids, d, cd, cb = oracle.beamsearch("SqL2", X, offs, adj, np.array([1], np.uint32), q, 1, 1,
                                       1.0, 10**6)
    assert ids[0, 0] == 50
    assert cd[0] == 1 + 50 and cb[0] == 50


def test_beam_maxvisits_cut(oracle):
    """Same walk cut by maxvisits = 10: the 11th inner evaluation (node 12) is pushed and
    ends the search (beamsearch.jl:61-63)."""
    X, offs, adj = _chain(70)
    q = (X[49] + 0.15)[None, :]
    ids, d, cd, cb = oracle.beamsearch("SqL2", X, offs, adj, np.array([1], np.uint32), q, 1, 1,
                                       1.0, 10)
    assert ids[0, 0] == 12
    assert cd[0] == 1 + 11 and cb[0] == 11


def test_beam_log2_fallback_without_hints(oracle):
    """No hints: ids 1:ceil(log2(1+n)) seed the result (beamsearch.jl:28-33)."""
    X, offs, adj = _chain(70)
    q = X[2:3]
    ids, d, cd, cb = oracle.beamsearch("SqL2", X, offs, adj, np.zeros(0, np.uint32), q, 3, 2,
                                       1.0, 10**6)
    # 7 fallback evaluations; the nearest (node 3) has both neighbours already seen
    assert cd[0] == 7
    np.testing.assert_array_equal(ids[0], [3, 2, 4])


def test_beam_small_n_is_brute_force(oracle):
    rng = np.random.default_rng(9)
    X = rng.random((40, 5), dtype=np.float32)
    Q = rng.random((15, 5), dtype=np.float32)
    offs = np.zeros(41, np.int64)
    ids, d, cd, cb = oracle.beamsearch("SqL2", X, offs, np.zeros(0, np.uint32),
                                       np.zeros(0, np.uint32), Q, 6, 4, 1.0, 10**6)
    gi, gd = oracle.exhaustive("SqL2", X, Q, 6)
    np.testing.assert_array_equal(ids, gi)
    assert (cd == 40).all() and (cb == 0).all()


# ---------------------------------------------------------------- Cosine
def test_cosine_closed_forms(oracle):
    """test/testdistances.jl:61-76: orthogonal vectors are at cosine distance 1, a vector
    (any length) from itself and from a positive multiple at 0."""
    e = np.array([[3.0, 0.0, 0.0]], np.float32)
    f = np.array([[0.0, 0.5, 0.0]], np.float32)
    E, F = oracle.unit_rows(e), oracle.unit_rows(f)
    assert oracle.dist("NormCosine", E[0], F[0]) == 1.0
    assert oracle.dist("NormCosine", E[0], E[0]) == 0.0
    assert oracle.dist("NormCosine", E[0], oracle.unit_rows(7 * e)[0]) == 0.0
    np.testing.assert_array_equal(oracle.unit_rows(np.zeros((1, 3), np.float32)), 0)


def test_cosine_within_fp32_bound_of_fp64(oracle):
    """The NormCosine distance of the normalised rows stays within 2 dim units of 2^-24 of
    the FP64 1 - a.b / (|a| |b|) (one rounding per normalised component and the final
    one), and a restatement that skipped the normalisation misses by far."""
    rng = np.random.default_rng(51)
    A = (rng.standard_normal((300, 96)) * rng.uniform(0.1, 10, (300, 1))).astype(np.float32)
    B = (rng.standard_normal((20, 96)) * 3).astype(np.float32)
    UA, UB = oracle.unit_rows(A), oracle.unit_rows(B)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    for j in range(len(B)):
        for i in range(len(A)):
            ref = 1 - A64[i] @ B64[j] / (np.linalg.norm(A64[i]) * np.linalg.norm(B64[j]))
            assert abs(oracle.dist("NormCosine", UA[i], UB[j]) - ref) <= 2 * 96 * 2.0 ** -24
    assert abs(oracle.dist("NormCosine", A[0], B[0]) -
               (1 - A64[0] @ B64[0] / (np.linalg.norm(A64[0]) * np.linalg.norm(B64[0])))) > 1e-3


# ---------------------------------------------------------------- BF16 form of the search
@pytest.mark.parametrize("metric", ["SqL2", "L2", "NormCosine"])
@pytest.mark.parametrize("dim", [8, 32, 96, 128])
def test_dist_bf16_exact_on_integer_rows(oracle, metric, dim):
    """Small integers are BF16 values and every FP32 step on them is exact, so the BF16
    search distance equals the canonical one bit for bit."""
    rng = np.random.default_rng(dim)
    X = rng.integers(-7, 8, (50, dim)).astype(np.float32)
    Q = rng.integers(-7, 8, (5, dim)).astype(np.float32)
    for q in Q:
        for x in X:
            assert oracle.dist_bf16(metric, x, q) == oracle.dist(metric, x, q)


@pytest.mark.parametrize("dim", [16, 96, 128, 256])
def test_dist_bf16_within_fp32_bound_of_rounded_rows(oracle, dim):
    """General rows: within the FP32 accumulation bound ((dim/16 + 4 + 3) units of 2^-24 of
    the sum of squares) of the FP64 distance between the BF16-rounded row and the FP32
    query; a restatement that skipped the BF16 rounding misses it by far."""
    rng = np.random.default_rng(dim + 1)
    X = rng.standard_normal((200, dim)).astype(np.float32)
    q = rng.standard_normal(dim).astype(np.float32)
    Xb = oracle.bf16_round(X)
    u = 2.0 ** -24
    worst_unrounded = 0.0
    for x, xb in zip(X, Xb):
        d = oracle.dist_bf16("SqL2", x, q)
        ref = ((xb.astype(np.float64) - q) ** 2).sum()
        assert abs(d - ref) <= (dim / 16 + 7) * u * ref
        ref_unrounded = ((x.astype(np.float64) - q) ** 2).sum()
        worst_unrounded = max(worst_unrounded, abs(d - ref_unrounded) / ((dim / 16 + 7) * u * ref))
    assert worst_unrounded > 10


def test_bf16_beamsearch_is_fp32_search_with_longer_queue_on_exact_rows(oracle):
    """On integer rows the BF16 search distance is exact, so the BF16 form with a result
    queue of max(k, rerank) entries is the FP32 search with k = rerank cut to its first
    k: same ids, distances and cost counters (the rerank keeps the queue's order)."""
    rng = np.random.default_rng(17)
    X = rng.integers(0, 12, (1500, 8)).astype(np.float32)
    Q = rng.integers(0, 12, (60, 8)).astype(np.float32)
    g = oracle.build_graph("SqL2", X, parallel_block=32, seed=4)
    args = ("SqL2", X, g["offs"], g["adj"], g["hints"], Q)
    for k, rr, bs in ((10, 14, 8), (5, 5, 16), (3, 20, 4)):
        bi, bd, bcd, bcb = oracle.beamsearch(*args, k, bs, 1.1, 10**6, bf16_rerank=rr)
        fi, fd, fcd, fcb = oracle.beamsearch(*args, max(k, rr), bs, 1.1, 10**6)
        np.testing.assert_array_equal(bi, fi[:, :k])
        np.testing.assert_array_equal(bd, fd[:, :k])
        np.testing.assert_array_equal(bcd, fcd)
        np.testing.assert_array_equal(bcb, fcb)


def test_bf16_beamsearch_returns_canonical_sorted_distances(oracle):
    """General rows: every returned distance is the canonical distance of its id, each
    column ascending, and the ids are distinct."""
    rng = np.random.default_rng(18)
    X = rng.random((1200, 16), dtype=np.float32)
    Q = rng.random((40, 16), dtype=np.float32)
    g = oracle.build_graph("SqL2", X, parallel_block=32, seed=6)
    ids, d, cd, cb = oracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 10, 8, 1.0,
                                       10**6, bf16_rerank=16)
    for j in range(len(Q)):
        assert len(set(ids[j].tolist())) == 10
        ex = [oracle.dist("SqL2", X[i - 1], Q[j]) for i in ids[j]]
        np.testing.assert_array_equal(d[j], np.array(ex, np.float32))
        assert np.all(np.diff(d[j]) >= 0)


@pytest.fixture(scope="module")
def built(oracle):
    rng = np.random.default_rng(11)
    X = rng.random((2000, 4), dtype=np.float32)
    g = oracle.build_graph("SqL2", X, parallel_block=32, seed=5)
    return X, g


def test_builder_graph_is_symmetric_without_loops(built):
    X, g = built
    offs, adj = g["offs"], g["adj"]
    n = len(X)
    edges = set()
    for i in range(n):
        for v in adj[offs[i]:offs[i + 1]]:
            assert v != i + 1
            edges.add((i + 1, int(v)))
    assert all((v, u) in edges for (u, v) in edges)


def test_builder_reaches_reference_recall_floor(oracle, built):
    """test/testsearchgraph.jl:48-75: n=2000, dim 4, 30 queries, recall >= 0.8."""
    X, g = built
    Q = np.random.default_rng(12).random((30, 4), dtype=np.float32)
    o = oracle.optimize_minrecall("SqL2", X, g["offs"], g["adj"], g["hints"], X[:64], 10, 0.9,
                                  cur_maxvisits=g["maxvisits"])
    ids, _, _, _ = oracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 10,
                                     o["bsize"], o["delta"], o["maxvisits"])
    gi, _ = oracle.exhaustive("SqL2", X, Q, 10)
    assert oracle.macrorecall(gi, ids) >= 0.8


def test_builder_hints_are_well_connected(built):
    X, g = built
    deg = np.diff(g["offs"])
    assert len(g["hints"]) > 0
    assert (deg[g["hints"].astype(np.int64) - 1] > 2).all()
    assert len(set(g["hints"].tolist())) == len(g["hints"])


def _hsp_forward(D, i, ksearch):
    """find_neighborhood! of object i (0-based) when every earlier object is a candidate
    (the n < 64 brute-force branch plus the interblock pushes, neighborhood.jl:25-45): the
    ksearch nearest earlier objects by (distance, id), then the proximal HSP filter
    (hsp.jl:103-119 with hsp_should_push's strict <, hsp.jl:8-33)."""
    prev = np.arange(i)
    order = prev[np.lexsort((prev, D[i, prev]))][:ksearch]
    kept = []
    for p in order:
        if all(D[p, h] >= D[i, p] for h in kept):
            kept.append(p)
    return [int(p) + 1 for p in kept]


@pytest.mark.parametrize("data", ["grid", "dim32", "axes"])
def test_builder_forward_lists_are_hsp_of_exact_candidates(oracle, data):
    """Objects 1..64 (blocks of 32) are inserted while the index holds fewer than 64
    objects, so their candidate lists are exact; each final list must start with the HSP
    filter of the ksearch = ceil(2 + log_1.3(block end)) nearest earlier objects
    (neighborhood.jl:5, insertions.jl:61), and carry every reverse link from a later
    object among the first 64 whose forward list names it.  Integer grid rows make exact
    distance ties (the filter's strict <, ties to the lower id); 32-dim rows keep long
    lists; unit basis vectors are all equally far apart, so the filter keeps every
    candidate and each list is exactly the neighbourhood size."""
    rng = np.random.default_rng(43)
    if data == "grid":
        X = rng.integers(0, 4, (300, 3)).astype(np.float32)
    elif data == "dim32":
        X = rng.random((300, 32), dtype=np.float32)
    else:
        X = np.concatenate([np.eye(64, dtype=np.float32), rng.random((236, 64), dtype=np.float32)])
    g = oracle.build_graph("SqL2", X, parallel_block=32, seed=5)
    n0 = 64
    D = oracle.dist_batch("SqL2", X[:n0], X[:n0]).reshape(n0, n0)
    fwd = {}
    for i in range(n0):
        ep = (i // 32 + 1) * 32
        ks = int(np.ceil(np.float32(2) + np.log(np.float32(ep)) / np.log(np.float32(1.3))))
        fwd[i] = _hsp_forward(D, i, ks)
    offs, adj = g["offs"], g["adj"]
    for i in range(n0):
        L = adj[offs[i]:offs[i + 1]].tolist()
        assert L[:len(fwd[i])] == fwd[i], i
        # the rest are reverse links, which only later objects make (connect_reverse_links!,
        # neighborhood.jl:52-89): among the first 64, exactly those naming i
        tail = L[len(fwd[i]):]
        assert all(x > i + 1 for x in tail), i
        assert {x for x in tail if x <= n0} == {j + 1 for j in range(n0) if (i + 1) in fwd[j]}, i
    assert sum(len(f) for f in fwd.values()) > 64


def test_optimizer_picks_the_minrecall_optimum_of_the_grid(oracle, built):
    """or_optimize_minrecall against an independent evaluation of every grid point
    (bsize 2:2:16 x delta 0.9:0.025:1.1, optbs.jl:28-33) with the pinned beam search and
    exhaustive gold: among points with recall >= r the least mean distance evaluations
    (first in grid order on ties), else the highest recall (opt.jl:281), and
    maxvisits = ceil(2 * visited.max) of the chosen point (optbs.jl:141-145), where the
    candidates ran with twice the current maxvisits (optbs.jl:153-156).  The visited
    statistics are per query: the reference's are per batch slot, which is per query when
    4 x nthreads >= the 64 sample queries."""
    X, g = built
    Q = np.random.default_rng(41).random((64, 4), dtype=np.float32)
    gold, _ = oracle.exhaustive("SqL2", X, Q, 10)
    for cur in (3000, 25):  # the second cuts searches at 2 * 25 evaluations
        _check_optimizer_grid(oracle, X, g, Q, gold, cur)


def _check_optimizer_grid(oracle, X, g, Q, gold, cur):
    pts = []
    for b in range(2, 17, 2):
        for di in range(9):
            d = np.float32((900.0 + 25.0 * di) / 1000.0)
            ids, _, cd, _ = oracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 10, b,
                                              d, 2 * cur)
            rec = np.mean([len(set(gold[j]) & set(ids[j][ids[j] != 0])) / 10 for j in range(64)])
            full = int(((ids != 0).sum(1) == 10).sum())
            act = cd[cd > 0]
            pts.append(dict(b=b, d=d, rec=rec, valid=full > 3, vmean=act.mean(), vmax=int(act.max())))
    for r in (0.9, 0.99, 1.01):
        o = oracle.optimize_minrecall("SqL2", X, g["offs"], g["adj"], g["hints"], Q, 10, r,
                                      cur_maxvisits=cur)
        valid = [p for p in pts if p["valid"]]
        ok = [p for p in valid if p["rec"] >= np.float32(r)]
        if ok:
            best = min(ok, key=lambda p: p["vmean"])   # min keeps the first in grid order
        else:
            best = max(valid, key=lambda p: p["rec"])  # max keeps the first in grid order
        assert (o["bsize"], np.float32(o["delta"])) == (best["b"], best["d"]), r
        assert o["maxvisits"] == int(np.ceil(2.0 * best["vmax"]))
        assert o["recall"] == pytest.approx(best["rec"])


# ---------------------------------------------------------------- fft, rerank, recall
def test_fft_invariants(oracle):
    rng = np.random.default_rng(13)
    X = rng.random((400, 3), dtype=np.float32)
    centers, nn, nnd, eps = oracle.fft("SqL2", X, 12, start=1)
    assert centers[0] == 1 and np.isinf(eps[0])
    assert np.all(np.diff(eps[1:]) <= 0)
    C = X[centers.astype(np.int64) - 1].astype(np.float64)
    d = ((X.astype(np.float64)[:, None, :] - C[None]) ** 2).sum(-1)
    np.testing.assert_allclose(nnd, d.min(1), rtol=1e-6, atol=1e-7)
    for c in range(1, 12):  # each new center was a farthest point at its round
        prev = d[:, :c].min(1)
        assert prev[centers[c] - 1] == pytest.approx(prev.max(), rel=1e-6)


def test_rerank_against_stable_sort(oracle):
    rng = np.random.default_rng(14)
    X = rng.random((300, 16), dtype=np.float32)
    Q = rng.random((20, 16), dtype=np.float32)
    ids = rng.integers(1, 301, (20, 12)).astype(np.uint32)
    ids[:, 9:] = 0
    gi, gd = oracle.rerank("SqL2", X, Q, ids, np.zeros((20, 12), np.float32))
    for j in range(20):
        d64 = ((X[ids[j, :9] - 1].astype(np.float64) - Q[j]) ** 2).sum(1)
        order = np.argsort(d64.astype(np.float32), kind="stable")
        np.testing.assert_array_equal(gi[j, :9], ids[j, :9][order])
        assert (gi[j, 9:] == 0).all()


def test_bf16_round_hand_cases(oracle):
    # exact values stay; halfway cases go to the even neighbour; overflow of the mantissa
    # carries into the exponent
    x = np.array([1.0, -2.5, 0.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20,
                  2 - 2 ** -9, 3.0e38], np.float32)
    want = np.array([1.0, -2.5, 0.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, 2.0, 2.996e38], np.float32)
    got = oracle.bf16_round(x)
    np.testing.assert_array_equal(got[:7], want[:7])
    assert abs(got[7] / 3.0e38 - 1) < 2 ** -8


def test_bf16_round_matches_torch(oracle):
    """torch's float32 -> bfloat16 conversion is round to nearest even; compared in a
    child process so that torch's own GPU runtime never shares a process with the HIP
    library."""
    import subprocess
    import sys
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        (rng.random(20000) * 1e-30).astype(np.float32),
                        rng.integers(-2 ** 20, 2 ** 20, 20000).astype(np.float32)])
    code = ("import sys, numpy as np, torch\n"
            "x = np.frombuffer(sys.stdin.buffer.read(), np.float32)\n"
            "y = torch.from_numpy(x.copy()).to(torch.bfloat16).to(torch.float32).numpy()\n"
            "sys.stdout.buffer.write(y.tobytes())\n")
    out = subprocess.run([sys.executable, "-c", code], input=x.tobytes(), capture_output=True,
                         check=True, timeout=300).stdout
    np.testing.assert_array_equal(oracle.bf16_round(x), np.frombuffer(out, np.float32))


def test_fp8_scale_exp_hand_cases(oracle):
    # the largest e with max|a| * 2^e <= 448 = 1.75 * 2^8
    assert oracle.fp8_scale_exp(np.array([448.0], np.float32)) == 0
    assert oracle.fp8_scale_exp(np.array([-449.0, 3.0], np.float32)) == -1
    assert oracle.fp8_scale_exp(np.array([1.0], np.float32)) == 8
    assert oracle.fp8_scale_exp(np.array([0.2, -1.75], np.float32)) == 8
    assert oracle.fp8_scale_exp(np.array([1.7500001], np.float32)) == 7
    assert oracle.fp8_scale_exp(np.array([0.5], np.float32)) == 9
    assert oracle.fp8_scale_exp(np.zeros(4, np.float32)) == 0
    assert oracle.fp8_scale_exp(np.array([1e-30], np.float32)) == 108
    assert oracle.fp8_scale_exp(np.array([1e-38], np.float32)) == 120   # clamped
    assert oracle.fp8_scale_exp(np.array([np.inf], np.float32)) == 0


def test_fp8_round_hand_cases(oracle):
    # exact values stay; halfway cases go to the even mantissa; subnormals have quantum
    # 2^-9; the largest value is 448; e scales and unscales exactly
    x = np.array([1.0, 448.0, -2.5, 1 + 2 ** -4, 1 + 3 * 2 ** -4, 1 + 2 ** -4 + 2 ** -20,
                  2 ** -10, 3 * 2 ** -10, 2 ** -6 - 2 ** -10, 2 - 2 ** -5, 0.0, 440.0], np.float32)
    want = [1.0, 448.0, -2.5, 1.0, 1.25, 1.125, 0.0, 2 ** -8, 2 ** -6, 2.0, 0.0, 448.0]
    np.testing.assert_array_equal(oracle.fp8_round(x, 0), want)
    np.testing.assert_array_equal(oracle.fp8_round(x / 8, 3), np.array(want) / 8)
    with pytest.raises(ValueError):
        oracle.fp8_round(np.array([449.0], np.float32), 0)


def test_fp8_round_matches_torch(oracle):
    """torch's float32 -> float8_e4m3fn conversion rounds to nearest even; compared in a
    child process so that torch's own GPU runtime never shares a process with the HIP
    library.  The values cover every binade of e4m3fn, its subnormals and exact ties."""
    import subprocess
    import sys
    rng = np.random.default_rng(23)
    ties = np.array([(2 * i + 1) * 2.0 ** -10 for i in range(8)] + [1 + (2 * i + 1) * 2 ** -4 for i in range(8)],
                    np.float32)
    x = np.concatenate([rng.uniform(-448, 448, 20000).astype(np.float32),
                        (rng.standard_normal(20000) * 0.05).astype(np.float32),
                        np.ldexp(rng.random(20000), rng.integers(-12, 9, 20000)).astype(np.float32),
                        ties, -ties])
    code = ("import sys, numpy as np, torch\n"
            "x = np.frombuffer(sys.stdin.buffer.read(), np.float32)\n"
            "y = torch.from_numpy(x.copy()).to(torch.float8_e4m3fn).to(torch.float32).numpy()\n"
            "sys.stdout.buffer.write(y.tobytes())\n")
    out = subprocess.run([sys.executable, "-c", code], input=x.tobytes(), capture_output=True,
                         check=True, timeout=300).stdout
    np.testing.assert_array_equal(oracle.fp8_round(x, 0), np.frombuffer(out, np.float32).astype(np.float64))


def _fp8_golden():
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_fp8_mfma
    return make_fp8_mfma, np.load(os.path.join(here, "fp8_mfma_cases.npz"), allow_pickle=False)


def _fp8_worst_units(oracle, mk, z, variant):
    """Largest |recorded - restated| over the recorded cases, in units of 2^-24 of the
    restated magnitude."""
    worst = 0.0
    for c in range(len(z["D"])):
        P = mk.decode(z["A"][c]).astype(np.float32)      # 16 rows x 128 k
        R = mk.decode(z["B"][c]).T.astype(np.float32)    # 16 columns x 128 k
        s = int(z["scale_a"][c]) + int(z["scale_b"][c]) - 254
        D, M = oracle.fp8_mfma(P, R, z["C"][c], 0, s, variant)
        worst = max(worst, float((np.abs(z["D"][c].astype(np.float64) - D) / (M * 2.0 ** -24)).max()))
    return worst


def test_fp8_mfma_restatement_matches_recorded_instruction(oracle):
    """or_fp8_mfma against what the FP8 matrix-core instruction returned on recorded
    random operands (tests/golden/fp8_mfma_cases.npz: every finite code, normal,
    positive, half-zero and subnormal-heavy operands, random scales, and the precision
    form's padded unit-vector rows with -|x|^2/2 start values): every output lies within
    3 units of 2^-24 of the magnitude of the restated truncated sum (the accumulation
    bound used in the parity tests is FP8_ACC_UNITS = 18), while a restatement without
    the truncation, with zeros in the anchor or with truncation toward -infinity misses
    by far more."""
    mk, z = _fp8_golden()
    assert _fp8_worst_units(oracle, mk, z, 0) <= 3.0
    assert oracle.FP8_ACC_UNITS >= 6 * 3.0
    for variant in (1, 2, 4):
        assert _fp8_worst_units(oracle, mk, z, variant) > 100.0, variant


def test_fp8_scores_pieces(oracle):
    """fp8_scores is or_fp8_mfma of fp8_operands started at -fma_norms / 2; the norms are
    the FP32 fma chain (within its rounding bound of the FP64 value); with operands whose
    products need no truncation the restated value is the exact dot product; products
    far below a group's largest are truncated toward zero."""
    rng = np.random.default_rng(24)
    X = rng.random((50, 20), dtype=np.float32)
    xn = oracle.fma_norms(X)
    x64 = (X.astype(np.float64) ** 2).sum(1)
    assert (np.abs(xn - x64) <= 21 * 2.0 ** -24 * x64).all()
    # powers of two with equal exponents in each group: every product is kept exactly
    P = np.ldexp(1.0, rng.integers(0, 3, (6, 128))).astype(np.float32)
    R = np.ldexp(1.0, -rng.integers(0, 3, (5, 128))).astype(np.float32)
    D, M = oracle.fp8_mfma(P, R, np.full(5, 0.25, np.float32), 1, 0)
    np.testing.assert_array_equal(D, 0.25 + P.astype(np.float64) @ R.T.astype(np.float64))
    np.testing.assert_array_equal(M, D)
    # 1 and 1.5 * 2^-13 (kept as 2^-13) in one group, -1.875 * 2^-14 (dropped) in it too,
    # and 2^-18 alone in the next group (kept)
    a = np.zeros((1, 128), np.float32)
    b = np.zeros((1, 128), np.float32)
    a[0, :3] = [1.0, 1.5 * 2 ** -4, -1.875 * 2 ** -5]
    b[0, :3] = [1.0, 2 ** -9, 2 ** -9]
    a[0, 8], b[0, 8] = 2 ** -9, 2 ** -9
    D, _ = oracle.fp8_mfma(a, b)
    assert D[0, 0] == 1.0 + 2 ** -13 + 2 ** -18


def test_prec_model_bounds_hold_in_fp32(oracle):
    """The BF16 precision form's error bound covers an FP32 evaluation of the same
    products in two different orders (a plausible accumulation of the matrix cores)."""
    rng = np.random.default_rng(22)
    for metric, X, Q in [
        ("NormCosine", *[(lambda a: a / np.linalg.norm(a, axis=1, keepdims=True))(
            rng.standard_normal((n, 96)).astype(np.float32)).astype(np.float32) for n in (300, 40)]),
        ("SqL2", rng.random((300, 64), dtype=np.float32), rng.random((40, 64), dtype=np.float32)),
    ]:
        A = oracle.prec_approx(metric, X, Q)
        B = oracle.prec_error_bound(metric, X, Q)
        Xb, Qb = oracle.bf16_round(X), oracle.bf16_round(Q)
        for order in (slice(None), slice(None, None, -1)):
            s = np.zeros((len(Q), len(X)), np.float32)
            for i in range(X.shape[1])[order]:
                s = (s + np.outer(Qb[:, i], Xb[:, i]).astype(np.float32)).astype(np.float32)
            if metric == "NormCosine":
                g = (np.float32(1) - s).astype(np.float64)
            else:
                xn = (X.astype(np.float32) ** 2).sum(1, dtype=np.float32)
                qn = (Q.astype(np.float32) ** 2).sum(1, dtype=np.float32)
                g = (qn[:, None] + xn[None, :] - np.float32(2) * s).astype(np.float64)
            assert (np.abs(g - A) <= B).all()
        # a dropped rounding step (FP32 operands instead of BF16) breaks the model
        s32 = Q.astype(np.float64) @ X.astype(np.float64).T
        g32 = 1 - s32 if metric == "NormCosine" else None
        if g32 is not None:
            assert (np.abs(g32 - A) > B).any()


def test_macrorecall(oracle):
    g = np.array([[1, 2, 3], [4, 5, 6]], np.uint32)
    assert oracle.macrorecall(g, g) == 1.0
    assert oracle.macrorecall(g, g + 10) == 0.0
    r = np.array([[3, 9, 1], [6, 7, 8]], np.uint32)
    assert oracle.macrorecall(g, r) == pytest.approx((2 / 3 + 1 / 3) / 2)


# ---------------------------------------------------------------- batch consumers
def _nearest_non_self(D, j):
    row = D[j].copy()
    row[j] = np.inf
    i = int(np.argmin(row))          # least distance, lowest id on ties
    return i, row[i]


@pytest.mark.parametrize("metric", ["SqL2", "L2", "NormCosine"])
def test_closestpair_against_all_pairs(oracle, metric):
    """closestpair: every pair's canonical distance (dist_batch) gives the closest pair,
    the lowest j first on ties; a planted exact duplicate is the pair found."""
    rng = np.random.default_rng(70)
    X = rng.random((300, 8), dtype=np.float32)
    if metric == "NormCosine":
        X /= np.linalg.norm(X, axis=1, keepdims=True)
    for planted in (False, True):
        if planted:
            X[16] = X[4]
        D = oracle.dist_batch(metric, X, X)
        best = (0, 0, np.inf)
        for j in range(len(X)):
            i, d = _nearest_non_self(D, j)
            if d < best[2]:
                best = (i + 1, j + 1, d)
        got = oracle.bichromatic_closestpair(metric, X, X, 8, samedata=True)
        assert got[:2] == best[:2] and got[2] == best[2]
        if planted:   # NormCosine of a row with itself is 1 - |x|^2, not always 0
            assert got[:2] == (17, 5) and got[2] == D[16, 4] and (metric == "NormCosine" or got[2] == 0)


def test_bichromatic_closestpair_against_all_pairs(oracle):
    rng = np.random.default_rng(71)
    A = rng.random((250, 6), dtype=np.float32)
    B = rng.random((120, 6), dtype=np.float32)
    D = oracle.dist_batch("SqL2", A, B)          # (|B|, |A|)
    j, i = np.unravel_index(np.argmin(D), D.shape)
    got = oracle.bichromatic_closestpair("SqL2", A, B, 4)
    assert got == (int(i) + 1, int(j) + 1, D[j, i])


def _greedy_net(D, eps):
    centers, nn, dd = [], np.zeros(len(D), np.int32), np.zeros(len(D), np.float32)
    for i in range(len(D)):
        if centers:
            row = D[i, centers]
            b = int(np.argmin(row))
            if row[b] <= eps:
                nn[i], dd[i] = centers[b] + 1, row[b]
                continue
        centers.append(i)
        nn[i], dd[i] = i + 1, 0.0
    return np.asarray(centers) + 1, nn, dd


@pytest.mark.parametrize("metric,eps", [("SqL2", 0.05), ("L2", 0.3), ("NormCosine", 0.02)])
def test_neardup_one_block_is_the_greedy_net(oracle, metric, eps):
    """With one block, neardup is the greedy eps-net in object order: an object joins
    the net unless a center is within eps, else it points to the first nearest center."""
    rng = np.random.default_rng(72)
    X = rng.random((400, 4), dtype=np.float32)
    if metric == "NormCosine":
        X /= np.linalg.norm(X, axis=1, keepdims=True)
    M, L, D, centers, cost = oracle.neardup(metric, X, eps, blocksize=400)
    c, nn, dd = _greedy_net(oracle.dist_batch(metric, X, X), np.float32(eps))
    np.testing.assert_array_equal(centers, c)
    np.testing.assert_array_equal(M, c)
    np.testing.assert_array_equal(L, nn)
    np.testing.assert_array_equal(D, dd)
    assert 1 < len(c) < 400


def test_neardup_cost_counts_net_searches_hand_case(oracle):
    """Blocks of 2 on a line (rows v * ones(4), SqL2, eps 1e-3): block 1 {0, 10} makes two
    centers and its filter evaluates object 2 against center 1 (1); block 2 {0.01, 20}
    searches the 2-center net (2 x 2 = 4): 0.01 is covered by 0, 20 becomes a center with no
    filter evaluation; block 3 {10.01, 20.01} searches the 3-center net (2 x 3 = 6), both
    covered.  costdists = 1 + 4 + 6 = 11 (sequential-exhaustive.jl:78, neardup.jl:240)."""
    v = np.array([0.0, 10.0, 0.01, 20.0, 10.01, 20.01], np.float32)
    X = (v[:, None] * np.ones((1, 4), np.float32)).astype(np.float32)
    M, L, D, cen, cost = oracle.neardup("SqL2", X, 1e-3, blocksize=2)
    assert cost == 11
    np.testing.assert_array_equal(cen, [1, 2, 4])
    np.testing.assert_array_equal(L, [1, 2, 1, 4, 2, 4])


def test_neardup_blocked_net_invariants(oracle):
    """Blocked runs: the centers are the objects that cover themselves, pairwise farther
    than eps; every other object lies within eps of its center, at the recorded distance;
    without filtering, every object of the first block is a center."""
    rng = np.random.default_rng(73)
    X = rng.random((900, 3), dtype=np.float32)
    eps = np.float32(0.01)
    for bs in (64, 200):
        M, L, D, centers, cost = oracle.neardup("SqL2", X, eps, blocksize=bs)
        np.testing.assert_array_equal(np.sort(centers), np.nonzero(L == np.arange(1, 901))[0] + 1)
        C = X[centers - 1]
        P = oracle.dist_batch("SqL2", C, C)
        np.fill_diagonal(P, np.inf)
        assert (P > eps).all()
        nc = L != np.arange(1, 901)
        assert (D[nc] <= eps).all()
        dd = np.array([oracle.dist("SqL2", X[i], X[L[i] - 1]) for i in np.nonzero(nc)[0]])
        np.testing.assert_array_equal(D[nc], dd)
        assert set(L[nc].tolist()) <= set(centers.tolist())
    M, L, D, centers, cost = oracle.neardup("SqL2", X, eps, blocksize=64, filterblocks=False)
    assert (centers[:64] == np.arange(1, 65)).all()


def test_hsp_queries_keep_drop_invariants(oracle):
    """hsp_queries, checked through properties of its output that fix it uniquely (by
    induction over the column order): the first candidate is kept; a kept candidate has
    no kept candidate before it in the column that is strictly nearer to it than its
    given distance, and a dropped one has; the kept fill the output as a stable sort by
    distance (column order on ties), then 0 / Inf.  Columns are given shuffled, with tied
    distances, and some end early (0)."""
    rng = np.random.default_rng(74)
    X = rng.random((500, 5), dtype=np.float32)
    Q = rng.random((40, 5), dtype=np.float32)
    ids, d = oracle.exhaustive("SqL2", X, Q, 16)
    for q in range(len(Q)):
        p = rng.permutation(16)
        ids[q], d[q] = ids[q, p], d[q, p]
    d[::3, 2] = d[::3, 5]
    ids[::7, 12:] = 0
    hi, hd = oracle.hsp_queries("SqL2", X, Q, ids, d)
    for q in range(len(Q)):
        L = int(np.argmax(ids[q] == 0)) if (ids[q] == 0).any() else 16
        P = oracle.dist_batch("SqL2", X[ids[q, :L] - 1], X[ids[q, :L] - 1])
        nk = int((hi[q] != 0).sum())
        pos = [int(np.nonzero(ids[q, :L] == v)[0][0]) for v in hi[q, :nk]]
        kept = set(pos)
        assert 0 in kept and 1 <= nk < L
        for c in range(L):
            blocked = any(P[c, h] < d[q, c] for h in kept if h < c)
            assert blocked == (c not in kept)
        order = sorted(kept, key=lambda c: (d[q, c], c))
        assert pos == order
        np.testing.assert_array_equal(hd[q, :nk], d[q, order])
        assert (hi[q, nk:] == 0).all() and np.isinf(hd[q, nk:]).all()


def test_hsp_queries_hand_case(oracle):
    """Points on a line, query at 0: a candidate at distance exactly equal to its
    distance from a kept one is kept (strict <), a farther-behind one is dropped."""
    X = np.zeros((4, 3), np.float32)
    X[:, 0] = [2.0, 1.0, 3.0, -1.0]                 # objects 1..4
    Q = np.zeros((1, 3), np.float32)
    ids = np.array([[1, 2, 3, 4]], np.uint32)        # given out of distance order
    d = np.array([[4.0, 1.0, 9.0, 1.0]], np.float32)
    hi, hd = oracle.hsp_queries("SqL2", X, Q, ids, d)
    # 1 kept; 2: d(2,1) = 1, not < 1, kept; 3: d(3,1) = 1 < 9, dropped;
    # 4: d(4,1) = 9, d(4,2) = 4, neither < 1, kept.  Sorted: 2 (1.0), 4 (1.0), 1 (4.0)
    np.testing.assert_array_equal(hi[0], [2, 4, 1, 0])
    np.testing.assert_array_equal(hd[0], np.array([1.0, 1.0, 4.0, np.inf], np.float32))


def test_staged_build_on_block_boundaries_is_one_build(oracle):
    """index! continued after append_items! starts a block at the first new object; when
    every stage ends on a block boundary the blocks, draws and callbacks are those of a
    single index!, so the graphs are identical (and a stage off a boundary changes them)."""
    from similaritysearch import datasets
    X = datasets.gmm(3000, 8, seed=3)
    one = oracle.build_graph("SqL2", X, parallel_block=64, seed=1)
    st = oracle.build_graph("SqL2", X, parallel_block=64, seed=1, stages=[640, 1280, 3000])
    for key in ("offs", "adj", "hints"):
        np.testing.assert_array_equal(one[key], st[key])
    assert (one["bsize"], one["delta"], one["maxvisits"]) == (st["bsize"], st["delta"], st["maxvisits"])
    off = oracle.build_graph("SqL2", X, parallel_block=64, seed=1, stages=[1000, 3000])
    assert len(off["adj"]) != len(one["adj"]) or not np.array_equal(off["adj"], one["adj"])


def test_neardup_graph_form_small_net_equals_exact(oracle):
    """While the net has fewer than 64 centers the beam search is brute force over it
    (beamsearch.jl:94-101) and charges the net's size per object, as the exact search does:
    the recall < 1 form then equals the exact form in every output."""
    rng = np.random.default_rng(74)
    C0 = rng.random((30, 4), dtype=np.float32)
    X = (C0[rng.integers(0, 30, 700)] + rng.normal(0, 1e-3, (700, 4))).astype(np.float32)
    eps = np.float32(0.01)
    a = oracle.neardup("SqL2", X, eps, blocksize=100)
    b = oracle.neardup("SqL2", X, eps, blocksize=100,
                       graph=dict(parallel_block=16, opt_minrecall=0.9, seed=2))
    assert len(a[3]) < 64
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, y)
    assert a[4] == b[4]


def test_neardup_graph_form_invariants(oracle):
    """With a larger net (approximate searches): centers cover themselves at distance 0,
    every other object lies within eps of a center, at the recorded distance."""
    rng = np.random.default_rng(75)
    X = rng.random((1500, 3), dtype=np.float32)
    eps = np.float32(0.004)
    M, L, D, centers, cost = oracle.neardup("SqL2", X, eps, blocksize=200,
                                            graph=dict(parallel_block=32, opt_minrecall=0.9, seed=4))
    assert len(centers) > 200
    np.testing.assert_array_equal(np.sort(centers), np.nonzero(L == np.arange(1, 1501))[0] + 1)
    other = L != np.arange(1, 1501)
    assert np.all(D[other] <= eps) and np.all(D[~other] == 0)
    assert set(L[other].tolist()) <= set(centers.tolist())
    d = oracle.dist_batch("SqL2", X, X)
    np.testing.assert_array_equal(D[other], d[np.nonzero(other)[0], L[other] - 1])
