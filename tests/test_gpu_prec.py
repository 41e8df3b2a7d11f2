"""GPU parity of the precision form of exhaustive k-NN and allknn: a candidate pass with
BF16 or FP8 (e4m3fn, power-of-two scaled) operands on the matrix cores keeps the ncand
best approximate candidates per query, then rerank! (src/rerank.jl:22-98) rescores them
with the canonical distance.

Pinned against the oracle:
- returned distances are the canonical FP32 distances of the returned ids (bit-exact)
  and results are in (distance, id) order;
- BF16: every approximate distance is within the FP32-accumulation bound of the FP64
  model of the BF16 operands (oracle.prec_approx / prec_error_bound), which fixes the
  operand rounding and the matrix-core data layout, and the candidate set is the ncand
  best under those approximate distances: no object outside it is provably better than
  one inside it;
- FP8: the candidate pass is restated bit for bit (oracle.fp8_scores: fp8_round with
  fp8_scale_exp, and the instruction's arithmetic pinned by tests/golden/
  fp8_mfma_cases.npz), so every approximate distance equals the restated one and the
  candidate set is exactly the ncand best by (score, id);
- the returned ids are the exact top k among the candidates;
- with ncand >= n every object is a candidate and the result equals the exact
  exhaustive search bit for bit;
- recall against the exact exhaustive search on random data.
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


def _dist(ss, metric):
    return {"SqL2": ss.SqL2, "L2": ss.L2, "NormCosine": ss.NormCosine}[metric]


def _check(oracle, metric, X, Q, k, ncand, ids, d, ci, cd, fmt="bf16", Qs=None):
    """The pinned properties for one call; ids/d are (m, k), ci/cd (m, ncand).  Qs is
    the whole query set of the call when Q is a subset of it (the FP8 query scale comes
    from every query)."""
    n = len(X)
    m = len(Q)
    D = oracle.dist_batch(metric, X, Q)            # canonical FP32 distances, (m, n)
    if fmt == "fp8":
        # the FP8 pass restated: e4m3fn operands, per-group truncation of the products,
        # and a bound for the FP32 accumulation
        A, B = oracle.fp8_approx(metric, X, Q, Qs)
    else:
        A = oracle.prec_approx(metric, X, Q)       # FP64 model of the approximate distances
        B = oracle.prec_error_bound(metric, X, Q)
    nc = min(ncand, n)
    for j in range(m):
        c = ci[j, :nc].astype(np.int64)
        assert (c >= 1).all() and (c <= n).all() and len(set(c.tolist())) == nc
        assert (ci[j, nc:] == 0).all() and np.isinf(cd[j, nc:]).all()
        c0 = c - 1
        # approximate distances within the accumulation bound of the model
        assert (np.abs(cd[j, :nc].astype(np.float64) - A[j, c0]) <= B[j, c0]).all(), j
        # no object outside the candidates is provably nearer than one inside
        if nc < n:
            out = np.ones(n, bool)
            out[c0] = False
            worst_in = (A[j, c0] - B[j, c0]).max()
            assert ((A[j, out] + B[j, out]) >= worst_in).all(), j
        # candidates in exact (distance, id) order
        key = np.lexsort((c0, D[j, c0]))
        np.testing.assert_array_equal(key, np.arange(nc))
        # the result is the exact top k among the candidates, with canonical distances
        kk = min(k, nc)
        np.testing.assert_array_equal(ids[j, :kk], ci[j, :kk])
        np.testing.assert_array_equal(d[j, :kk], D[j, c0[:kk]])
        assert (ids[j, kk:] == 0).all() and np.isinf(d[j, kk:]).all()


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
@pytest.mark.parametrize("metric,dim,n,m,k,ncand", [
    ("SqL2", 32, 200, 300, 10, 256),
    ("NormCosine", 96, 150, 70, 32, 128),
    ("SqL2", 16, 9, 5, 12, 16),
    ("SqL2", 1, 40, 7, 5, 40),
    ("SqL2", 24, 300, 0, 6, 64),   # m = 0: eight all-zero queries
])
def test_every_object_a_candidate_equals_exact(oracle, metric, dim, n, m, k, ncand, fmt):
    ss = _ss()
    X = _unit(n, dim, 1) if metric == "NormCosine" else _uniform(n, dim, 1)
    Q = _unit(m, dim, 2) if metric == "NormCosine" else _uniform(m, dim, 2)
    if m == 0:
        Q = np.zeros((8, dim), np.float32)
    E = ss.ExhaustiveSearch(_dist(ss, metric), ss.MatrixDatabase.from_rows(X))
    ids, d, ci, cd = ss.exhaustive_candidates(E, k, ncand, ss.MatrixDatabase.from_rows(Q), fmt)
    gi, gd = oracle.exhaustive(metric, X, Q, k)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    _check(oracle, metric, X, Q, k, ncand, ids.T, d.T, ci.T, cd.T, fmt)


# recall@k against the exact search: BF16 operands keep the top k within the candidates
# on these shapes; FP8 (3 mantissa bits) needs the margin of ncand over k
_MIN_RECALL = {"bf16": 0.99, "fp8": 0.95}


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
@pytest.mark.parametrize("metric,dim,n,m,k,ncand,slices", [
    ("NormCosine", 96, 6000, 300, 32, 128, 0),
    ("NormCosine", 96, 6000, 300, 32, 128, 3),
    ("SqL2", 64, 5000, 257, 10, 64, 0),
    ("SqL2", 128, 4000, 100, 10, 40, 5),
    ("SqL2", 20, 3000, 64, 1, 32, 2),
    # lists of 1024 entries (ncand > 128), compacted several times per slice
    ("NormCosine", 96, 6000, 300, 32, 200, 0),
    ("SqL2", 128, 4000, 100, 10, 160, 2),
])
def test_candidates_and_rerank(oracle, metric, dim, n, m, k, ncand, slices, fmt):
    ss = _ss()
    from similaritysearch import _lib
    X = _unit(n, dim, 3) if metric == "NormCosine" else _uniform(n, dim, 3)
    Q = _unit(m, dim, 4) if metric == "NormCosine" else _uniform(m, dim, 4)
    E = ss.ExhaustiveSearch(_dist(ss, metric), ss.MatrixDatabase.from_rows(X))
    _lib.check(_lib.load().ssg_set_option(E._handle(), _lib.SSG_OPT_PREC_SLICES, slices), "option")
    ids, d, ci, cd = ss.exhaustive_candidates(E, k, ncand, ss.MatrixDatabase.from_rows(Q), fmt)
    _check(oracle, metric, X, Q, k, ncand, ids.T, d.T, ci.T, cd.T, fmt)
    gi, _ = oracle.exhaustive(metric, X, Q, k)
    rec = oracle.macrorecall(gi, ids.T)
    print(f"recall {fmt} {metric} dim {dim} n {n} k {k} ncand {ncand}: {rec:.4f}")
    assert rec >= _MIN_RECALL[fmt], rec


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
def test_l2_precision_form_follows_sqeuclidean(oracle, fmt):
    """L2 scores the candidates as SqL2 does (the same -|x|^2/2 start value), so its
    candidate set equals SqL2's, its approximate distances are their square roots (FP32
    sqrt of the reported SqL2 value, clamped at 0; within one ulp), and its results are
    the canonical L2 distances of the exact top k among the candidates."""
    ss = _ss()
    X = _uniform(5000, 48, 50)
    Q = _uniform(120, 48, 51)
    E2 = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    E1 = ss.ExhaustiveSearch(ss.L2, ss.MatrixDatabase.from_rows(X))
    i2, d2, c2, a2 = [t.T for t in ss.exhaustive_candidates(E2, 10, 48, ss.MatrixDatabase.from_rows(Q), fmt)]
    i1, d1, c1, a1 = [t.T for t in ss.exhaustive_candidates(E1, 10, 48, ss.MatrixDatabase.from_rows(Q), fmt)]
    for j in range(len(Q)):
        assert set(c1[j].tolist()) == set(c2[j].tolist()), j
        o1, o2 = np.argsort(c1[j]), np.argsort(c2[j])
        # within one ulp: the device's FP32 square root against numpy's
        np.testing.assert_array_max_ulp(a1[j][o1], np.sqrt(np.maximum(a2[j][o2], np.float32(0))), maxulp=1)
    D = oracle.dist_batch("L2", X, Q)
    gi, _ = oracle.exhaustive("L2", X, Q, 10)
    for j in range(len(Q)):
        np.testing.assert_array_equal(d1[j], D[j, i1[j].astype(np.int64) - 1])
        # the exact top 10 among the candidates, in (distance, id) order
        cd = D[j, c1[j].astype(np.int64) - 1]
        order = np.lexsort((c1[j], cd))[:10]
        np.testing.assert_array_equal(i1[j], c1[j][order])
    assert oracle.macrorecall(gi, i1) >= _MIN_RECALL[fmt]


def test_fp8_power_of_two_scaling_is_exact():
    """Scaling the data by 2^s leaves the FP8 operands unchanged (the scale exponents
    absorb it), so the candidates are the same and every approximate SqL2 distance scales
    by exactly 2^(2 s): the matrix-core scale operands and the -|x|^2/2 start value are
    in the same units."""
    ss = _ss()
    X = _uniform(3000, 40, 30)
    Q = _uniform(70, 40, 31)
    outs = []
    for s in (0, -9, 11):
        f = np.float32(2.0 ** s)
        E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X * f))
        outs.append(ss.exhaustive_candidates(E, 10, 64, ss.MatrixDatabase.from_rows(Q * f), "fp8"))
    for s, o in zip((-9, 11), outs[1:]):
        np.testing.assert_array_equal(o[2], outs[0][2])
        np.testing.assert_array_equal(o[0], outs[0][0])
        np.testing.assert_array_equal(o[3], outs[0][3] * np.float32(2.0 ** (2 * s)))


def test_formats_share_one_handle(oracle):
    """Switching the format on one index re-encodes its rows: bf16, fp8, bf16 again and
    fp8 again give the results of fresh indexes."""
    ss = _ss()
    X = _unit(4000, 64, 32)
    Q = _unit(130, 64, 33)
    E = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
    seq = [ss.exhaustive_candidates(E, 8, 32, ss.MatrixDatabase.from_rows(Q), f)
           for f in ("bf16", "fp8", "bf16", "fp8")]
    for f, o in zip(("bf16", "fp8", "bf16", "fp8"), seq):
        F = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
        r = ss.exhaustive_candidates(F, 8, 32, ss.MatrixDatabase.from_rows(Q), f)
        for a, b in zip(o, r):
            np.testing.assert_array_equal(a, b)
    _check(oracle, "NormCosine", X, Q, 8, 32, *[t.T for t in seq[1]], fmt="fp8")


def test_slice_count_does_not_change_results():
    ss = _ss()
    from similaritysearch import _lib
    X = _unit(7000, 96, 5)
    Q = _unit(260, 96, 6)
    E = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
    outs = []
    for s in (1, 2, 7, 0):
        _lib.check(_lib.load().ssg_set_option(E._handle(), _lib.SSG_OPT_PREC_SLICES, s), "option")
        outs.append(ss.exhaustive_candidates(E, 16, 64, ss.MatrixDatabase.from_rows(Q)))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
def test_seeded_sample_pass_does_not_change_results(oracle, fmt):
    """On slices of 256 Ki rows or more, a seed pass over every 16th tile of the sample
    starts the sample pass's thresholds.  Results equal those with sampling off, for the
    search and for allknn, and the candidates pass the oracle's checks."""
    ss = _ss()
    from similaritysearch import _lib
    L = _lib.load()
    X = _unit(300_000, 32, 60)
    Q = _unit(300, 32, 61)
    E = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
    h = E._handle()
    _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SLICES, 1), "option")
    outs = []
    for t in (0, -1):  # 0: sampled thresholds with the seed pass (1 slice of 300,000 rows)
        _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SAMPLE, t), "option")
        outs.append(ss.exhaustive_candidates(E, 16, 64, ss.MatrixDatabase.from_rows(Q), fmt))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    ids, d, ci, cd = outs[0]
    _check(oracle, "NormCosine", X, Q[:12], 16, 64, ids.T[:12], d.T[:12], ci.T[:12], cd.T[:12], fmt, Q)
    _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SAMPLE, 0), "option")


@pytest.mark.parametrize("fmt,slices", [("bf16", 1), ("bf16", 3), ("fp8", 2)])
def test_sampled_thresholds_do_not_change_results(oracle, slices, fmt):
    """The sampled starting thresholds only drop rows that cannot be candidates, and
    queries the rerank stage cannot certify are searched again: results are identical
    with the sample pass off, on (T = 24) and on with a threshold so high (T = 2) that
    most queries fail the certificate and go through the redo."""
    ss = _ss()
    from similaritysearch import _lib
    L = _lib.load()
    X = _unit(20000, 96, 15)
    Q = _unit(600, 96, 16)
    E = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
    h = E._handle()
    _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SLICES, slices), "option")
    outs = []
    for t in (-1, 24, 2):
        _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SAMPLE, t), "option")
        outs.append(ss.exhaustive_candidates(E, 32, 128, ss.MatrixDatabase.from_rows(Q), fmt))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            np.testing.assert_array_equal(a, b)
    ids, d, ci, cd = outs[2]
    _check(oracle, "NormCosine", X, Q[:40], 32, 128, ids.T[:40], d.T[:40], ci.T[:40], cd.T[:40], fmt, Q)
    # allknn form through the same path
    _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SAMPLE, 2), "option")
    a2 = ss.exhaustive_candidates(E, 16, 64, None, fmt)
    _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SAMPLE, -1), "option")
    a1 = ss.exhaustive_candidates(E, 16, 64, None, fmt)
    for a, b in zip(a1, a2):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
def test_seeded_redo_and_last_resort(oracle, fmt):
    """Slices of at least 256 Ki rows get seed thresholds.  With T = 2 the main pass cannot
    certify most queries; they are redone from the largest of their slices' seeds (one
    float step lower), and a query the redo cannot certify either is searched once more
    from no threshold: 100 copies of it among the rows, 8 of them in rows 0-7 (the first
    tile, in the seed sample), so its seed is its own score and the redo keeps only the
    100 copies of the 128 candidates it needs.  The results equal the search without
    sampling and hold the pinned properties of the candidate pass."""
    ss = _ss()
    from similaritysearch import _lib
    L = _lib.load()
    rng = np.random.default_rng(21)
    n, dim = 270000, 16
    X = rng.random((n, dim), dtype=np.float32)
    v = rng.random(dim, dtype=np.float32)
    pos = np.sort(np.concatenate([np.arange(8), 8 + rng.choice(n - 8, 92, replace=False)]))
    X[pos] = v
    Q = np.concatenate([v[None, :], rng.random((299, dim), dtype=np.float32)]).astype(np.float32)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    h = E._handle()
    _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SLICES, 1), "option")
    outs = []
    for t in (-1, 2):
        _lib.check(L.ssg_set_option(h, _lib.SSG_OPT_PREC_SAMPLE, t), "option")
        outs.append(ss.exhaustive_candidates(E, 10, 128, ss.MatrixDatabase.from_rows(Q), fmt))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
    ids, d, ci, cd = (o.T for o in outs[1])
    np.testing.assert_array_equal(ids[0], pos[:10] + 1)
    _check(oracle, "SqL2", X, Q[:6], 10, 128, ids[:6], d[:6], ci[:6], cd[:6], fmt, Q)


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
def test_ties_prefer_lower_ids(oracle, fmt):
    """300 copies of one vector among random rows: all have the same approximate and
    exact distance to it, so the candidate pass must keep the lowest ids (a compaction
    with a tie at its threshold) and the result is the k lowest copies."""
    ss = _ss()
    rng = np.random.default_rng(7)
    X = rng.random((5000, 32), dtype=np.float32)
    v = rng.random(32, dtype=np.float32)
    pos = np.sort(rng.choice(5000, 300, replace=False))
    X[pos] = v
    Q = np.stack([v, v + np.float32(1e-3)]).astype(np.float32)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    from similaritysearch import _lib
    # one slice: the ties are resolved by the compactions of one list
    _lib.check(_lib.load().ssg_set_option(E._handle(), _lib.SSG_OPT_PREC_SLICES, 1), "option")
    ids, d, ci, cd = ss.exhaustive_candidates(E, 10, 64, ss.MatrixDatabase.from_rows(Q), fmt)
    np.testing.assert_array_equal(ci.T[0], pos[:64] + 1)
    np.testing.assert_array_equal(ids.T[0], pos[:10] + 1)
    assert (d.T[0] == 0).all()
    gi, gd = oracle.exhaustive("SqL2", X, Q, 10)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)


@pytest.mark.parametrize("fmt,min_recall", [("bf16", 0.999), ("fp8", 0.99)])
def test_allknn_precision_form(oracle, fmt, min_recall):
    """allknn with the BF16 or FP8 candidate pass (1M x 96 unit vectors, k 32, ncand 128,
    at reduced n): self first, exact distances, recall against the exact allknn."""
    ss = _ss()
    X = _unit(8000, 96, 9)
    E = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
    ctx = ss.GenericContext(precision=fmt)
    ids, d = ss.allknn(E, ctx, 32)
    ids, d = ids.T, d.T
    assert (ids[:, 0] == np.arange(1, 8001)).all()
    sub = np.arange(0, 8000, 40)
    gi, gd = oracle.exhaustive("NormCosine", X, X[sub], 32)
    D = oracle.dist_batch("NormCosine", X, X[sub])
    for r, j in enumerate(sub):
        np.testing.assert_array_equal(d[j], D[r, ids[j].astype(np.int64) - 1])
    rec = oracle.macrorecall(gi, ids[sub])
    print(f"allknn recall {fmt}: {rec:.4f}")
    assert rec >= min_recall, rec
    # the candidate form of allknn agrees with the plain one
    i2, d2, ci, cd = ss.exhaustive_candidates(E, 32, 128, None, fmt)
    np.testing.assert_array_equal(i2.T, ids)
    np.testing.assert_array_equal(d2.T, d)
    _check(oracle, "NormCosine", X, X[sub[:25]], 32, 128, ids[sub[:25]], d[sub[:25]],
           ci.T[sub[:25]], cd.T[sub[:25]], fmt, X)


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
def test_searchbatch_precision_context(oracle, fmt):
    ss = _ss()
    X = _uniform(3000, 48, 10)
    Q = _uniform(90, 48, 11)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    ids, d = ss.searchbatch(E, ss.GenericContext(precision=fmt, ncand=64),
                            ss.MatrixDatabase.from_rows(Q), 8)
    gi, gd = oracle.exhaustive("SqL2", X, Q, 8)
    assert oracle.macrorecall(gi, ids.T) >= _MIN_RECALL[fmt]
    D = oracle.dist_batch("SqL2", X, Q)
    for j in range(90):
        np.testing.assert_array_equal(d.T[j], D[j, ids.T[j].astype(np.int64) - 1])


def test_precision_form_limits():
    ss = _ss()
    from similaritysearch import _lib
    L = _lib.load()
    X = _uniform(500, 16, 12)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    h = E._handle()
    Q = _uniform(4, 16, 13)
    ids = np.zeros((4, 300), np.uint32)
    d = np.zeros((4, 300), np.float32)
    assert L.ssg_exhaustive_knn_prec(h, Q.ctypes.data, 4, 10, 1, 300, ids.ctypes.data,
                                     d.ctypes.data, None, None) == -5
    assert L.ssg_exhaustive_knn_prec(h, Q.ctypes.data, 4, 10, 1, 8, ids.ctypes.data,
                                     d.ctypes.data, None, None) == -5
    assert L.ssg_exhaustive_knn_prec(h, Q.ctypes.data, 4, 10, 0, 64, ids.ctypes.data,
                                     d.ctypes.data, None, None) == -1
    Xw = _uniform(100, 190, 14)
    W = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(Xw))
    Qw = _uniform(2, 190, 15)
    assert L.ssg_exhaustive_knn_prec(W._handle(), Qw.ctypes.data, 2, 4, 1, 64, ids.ctypes.data,
                                     d.ctypes.data, None, None) == -5
