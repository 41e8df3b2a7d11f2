"""Parity at dimension 128, the dimension of the 1M x 128 search workload: the kernels
instantiated for it (beam_kernel<8, *>, beam_bf16_kernel<8, *>, gemm_f16_kernel<128>)
against the oracle on a 20,000 x 128 Gaussian mixture with a graph built on the GPU.

- FP32 beam search (beamsearch.jl:44-110): ids, distances, costdists and costblocks
  bit-exact against or_beamsearch.
- BF16 form: bit-exact against or_beamsearch_bf16 (the same search on BF16 rows with FP32
  arithmetic, then rerank! with the canonical distance); its recall is within 0.002 of
  the FP32 search's.
- Mixed per-query k / bsize batch, longest-first order and HIP-graph replay.
- The 4-round gather kernels and other launch shapes (ssg_set_launch) give the same bits,
  among them LDS visited tables of every size: none, a tiny one whose buckets fill at once
  (most marks in HBM, erased tags after the maxvisits cut), the automatic one and the
  largest;
  the launch autotuner (ssg_optimize_launch) installs a shape and reports the recall the
  search reaches.
- Learned form at D = 128, H = 256, L = 4096: one training step's gradients against the
  FP64 oracle (the forward output layer runs on the 128-tile GEMM), inference against a
  NumPy forward, and the planned search bit-exact against the oracle with the same hints
  and beam sizes.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, DIM = 20000, 128


@pytest.fixture(scope="module")
def g128():
    import similaritysearch as ss
    from similaritysearch import datasets
    X = datasets.gmm(N, DIM, seed=201)
    Q = datasets.gmm(500, DIM, seed=202)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=256, seed=5)
    offs = np.concatenate([[0], G.adj.offset]).astype(np.int64)
    return X, Q, G, offs


@pytest.mark.parametrize("bsize,delta,beamcap,maxvisits", [
    (12, 1.0, 32, 10**6),
    (16, 1.1, 32, 10**6),
    (32, 1.1, 32, 10**6),
    (16, 1.1, 256, 10**6),
    (32, 1.0, 256, 10**6),
    (100, 1.05, 256, 10**6),
    (16, 1.1, 32, 400),
])
def test_fp32_beamsearch_dim128_bit_exact(oracle, g128, bsize, delta, beamcap, maxvisits):
    import similaritysearch as ss
    X, Q, G, offs = g128
    G.algo = ss.BeamSearch(bsize, delta, maxvisits)
    ctx = ss.SearchGraphContext(beamcap=beamcap)
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
    gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 10, bsize,
                                         delta, maxvisits, beamcap=beamcap, nthreads=8)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    np.testing.assert_array_equal(ctx.costdists, gcd)
    np.testing.assert_array_equal(ctx.costblocks, gcb)


@pytest.mark.parametrize("k,bsize,beamcap,maxvisits", [
    (65, 12, 32, 10**6),
    (100, 16, 32, 10**6),
    (256, 64, 256, 10**6),
    (128, 32, 256, 700),
])
def test_fp32_beamsearch_large_k_bit_exact(oracle, g128, k, bsize, beamcap, maxvisits):
    """k beyond one register (the result queue in four: k <= 256), as KnnSorted(k) of any
    size in the reference."""
    import similaritysearch as ss
    X, Q, G, offs = g128
    G.algo = ss.BeamSearch(bsize, 1.1, maxvisits)
    ctx = ss.SearchGraphContext(beamcap=beamcap)
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), k)
    gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, k, bsize,
                                         1.1, maxvisits, beamcap=beamcap, nthreads=8)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    np.testing.assert_array_equal(ctx.costdists, gcd)
    np.testing.assert_array_equal(ctx.costblocks, gcb)


def test_large_k_queue_vector_and_allknn(oracle, g128, allknn128):
    """Per-query k up to 200 in the queue-vector form, and allknn with k = 100."""
    import similaritysearch as ss
    X, Q, G, offs = g128
    rng = np.random.default_rng(231)
    ks = rng.choice([1, 10, 100, 200], len(Q)).astype(np.int32)
    G.algo = ss.BeamSearch(16, 1.1, 10**6)
    ctx = ss.SearchGraphContext()
    res = ss.searchbatch_knns(G, ctx, ss.MatrixDatabase.from_rows(Q), ks)
    gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 200, 16,
                                         1.1, 10**6, kq=ks, nthreads=8)
    for j, (ij, dj) in enumerate(res):
        np.testing.assert_array_equal(ij, gi[j, :ks[j]])
        np.testing.assert_array_equal(dj, gd[j, :ks[j]])
    np.testing.assert_array_equal(ctx.costdists, gcd)
    Ga, _, _, Xa, ga = allknn128
    ids, d = ss.allknn(Ga, ss.SearchGraphContext(), 100)
    ai, ad = oracle.allknn_graph("SqL2", Xa, ga["offs"], ga["adj"], 100, 16, 1.1, 10**6, nthreads=8)
    np.testing.assert_array_equal(ids.T, ai)
    np.testing.assert_array_equal(d.T, ad)


def test_large_k_limits(g128):
    """k > 64 is refused in the BF16 form, and k > 256 everywhere."""
    import similaritysearch as ss
    from similaritysearch import _lib
    X, Q, G, offs = g128
    G.algo = ss.BeamSearch(16, 1.1, 10**6)
    with pytest.raises(_lib.SSGError):
        ss.searchbatch(G, ss.SearchGraphContext(precision="bf16", rerank=14),
                       ss.MatrixDatabase.from_rows(Q[:8]), 100)
    with pytest.raises(_lib.SSGError):
        ss.searchbatch(G, ss.SearchGraphContext(), ss.MatrixDatabase.from_rows(Q[:8]), 257)


@pytest.mark.parametrize("bsize,beamcap,rerank,k", [
    (16, 32, 14, 10),
    (32, 32, 14, 10),
    (100, 128, 20, 10),
    (64, 256, 14, 10),
    (16, 32, 10, 10),
    (24, 32, 40, 32),
])
def test_bf16_beamsearch_dim128_bit_exact(oracle, g128, bsize, beamcap, rerank, k):
    import similaritysearch as ss
    X, Q, G, offs = g128
    G.algo = ss.BeamSearch(bsize, 1.1, 10**6)
    ctx = ss.SearchGraphContext(beamcap=beamcap, precision="bf16", rerank=rerank)
    ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), k)
    gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, k, bsize,
                                         1.1, 10**6, beamcap=beamcap, nthreads=8,
                                         bf16_rerank=rerank)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(d.T, gd)
    np.testing.assert_array_equal(ctx.costdists, gcd)
    np.testing.assert_array_equal(ctx.costblocks, gcb)


@pytest.mark.parametrize("bsize", [16, 32])
def test_bf16_recall_within_0002_of_fp32(oracle, g128, bsize):
    """The BF16 form with a rerank of 14 finds the exact top-10 as often as the FP32 search
    (recall@10 against exhaustive gold within 0.002 on 500 queries)."""
    import similaritysearch as ss
    X, Q, G, offs = g128
    G.algo = ss.BeamSearch(bsize, 1.1, 10**6)
    Qdb = ss.MatrixDatabase.from_rows(Q)
    i32, _ = ss.searchbatch(G, ss.SearchGraphContext(), Qdb, 10)
    i16, _ = ss.searchbatch(G, ss.SearchGraphContext(precision="bf16", rerank=14), Qdb, 10)
    gold, _ = oracle.exhaustive("SqL2", X, Q, 10, nthreads=8)
    r32 = oracle.macrorecall(gold, i32.T)
    r16 = oracle.macrorecall(gold, i16.T)
    assert r16 >= r32 - 0.002, (r16, r32)


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_mixed_batch_dim128_under_graph_replay(oracle, g128, prec):
    """200k-query form at small m: per-query k ~ {1, 10, 50}, bsize ~ {16..256}, one
    persistent launch, longest-first order, recorded once and replayed three times."""
    import similaritysearch as ss
    from similaritysearch import _lib, datasets
    from similaritysearch.device import DeviceArray, Stream
    X, _, G, offs = g128
    rng = np.random.default_rng(203)
    Q = datasets.gmm(900, DIM, seed=204)
    m = len(Q)
    kq = rng.choice([1, 10, 50], m).astype(np.int32)
    bq = rng.choice([16, 32, 64, 128, 256], m).astype(np.int32)
    rr = 20 if prec == "bf16" else None
    oi, od, ocd, ocb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 10, 16, 1.1,
                                         10**6, beamcap=256, kq=kq, bq=bq, nthreads=8,
                                         bf16_rerank=rr)
    L = _lib.load()
    h = G._handle()
    st = Stream()
    dQ, dk, db = DeviceArray.from_numpy(Q), DeviceArray.from_numpy(kq), DeviceArray.from_numpy(bq)
    dI, dD = DeviceArray((m, 50), np.uint32), DeviceArray((m, 50), np.float32)
    dcd, dcb, dO = DeviceArray((m,), np.int64), DeviceArray((m,), np.int64), DeviceArray((m,), np.uint32)
    bp = ss.BeamParams(16, 1.1, 10**6, 256, 1 if prec == "bf16" else 0, rr or 0, 1)
    _lib.check(L.ssg_order_by_bsize_dev(h, db.ptr, m, dO.ptr, st.handle), "order")
    qo = ss.QueryOpts(dk.ptr, db.ptr, None, None, 0, dO.ptr)

    def run():
        _lib.check(L.ssg_beamsearch_dev(h, dQ.ptr, m, 50, 50, C.byref(bp), C.byref(qo), dI.ptr,
                                        dD.ptr, dcd.ptr, dcb.ptr, st.handle), "searchbatch!")

    def check():
        np.testing.assert_array_equal(dI.numpy(), oi)
        np.testing.assert_array_equal(dD.numpy(), od)
        np.testing.assert_array_equal(dcd.numpy(), ocd)
        np.testing.assert_array_equal(dcb.numpy(), ocb)

    run()
    st.synchronize()
    check()
    _lib.check(L.ssg_capture_begin(st.handle), "capture")
    try:
        run()
    finally:
        g = C.c_void_p()
        rc = L.ssg_capture_end(st.handle, C.byref(g))
    _lib.check(rc, "capture end")
    try:
        for _ in range(3):
            dI.upload(np.zeros((m, 50), np.uint32))
            dD.upload(np.zeros((m, 50), np.float32))
            dcd.upload(np.zeros(m, np.int64))
            _lib.check(L.ssg_graph_launch(g, st.handle), "replay")
            st.synchronize()
            check()
    finally:
        L.ssg_graph_destroy(g)


@pytest.mark.parametrize("prec", ["f32", "bf16"])
@pytest.mark.parametrize("shape", [(8, 4, -1), (16, 4, 0), (64, 8, -1), (4, 8, 32), (24, 10, -1),
                                   (16, 8, 2048), (16, 10, 32)])
def test_launch_shapes_are_bit_exact(oracle, g128, prec, shape):
    import similaritysearch as ss
    from similaritysearch import _lib
    X, Q, G, offs = g128
    L = _lib.load()
    h = G._handle()
    _lib.check(L.ssg_set_launch(h, C.byref(_lib.LaunchParams(*shape))), "set launch")
    try:
        G.algo = ss.BeamSearch(24, 1.1, 10**6)
        rr = 14 if prec == "bf16" else None
        ctx = ss.SearchGraphContext(precision=prec, rerank=rr or 0)
        ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
        gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 10, 24,
                                             1.1, 10**6, nthreads=8, bf16_rerank=rr)
        np.testing.assert_array_equal(ids.T, gi)
        np.testing.assert_array_equal(d.T, gd)
        np.testing.assert_array_equal(ctx.costdists, gcd)
        np.testing.assert_array_equal(ctx.costblocks, gcb)
    finally:
        _lib.check(L.ssg_set_launch(h, C.byref(_lib.LaunchParams(32, 8, -1))), "set launch")


@pytest.mark.parametrize("prec", ["f32", "bf16"])
@pytest.mark.parametrize("beamcap", [128, 256])
def test_wide_beams_4_rounds_bit_exact(oracle, g128, prec, beamcap):
    """The 4-round kernels of the 2- and 4-register beams (fewer registers, more waves
    per SIMD) give the oracle's bits."""
    import similaritysearch as ss
    from similaritysearch import _lib
    X, Q, G, offs = g128
    L = _lib.load()
    h = G._handle()
    _lib.check(L.ssg_set_launch(h, C.byref(_lib.LaunchParams(32, 4, -1))), "set launch")
    try:
        G.algo = ss.BeamSearch(100, 1.05, 10**6)
        rr = 14 if prec == "bf16" else None
        ctx = ss.SearchGraphContext(beamcap=beamcap, precision=prec, rerank=rr or 0)
        ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
        gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 10, 100,
                                             1.05, 10**6, beamcap=beamcap, nthreads=8, bf16_rerank=rr)
        np.testing.assert_array_equal(ids.T, gi)
        np.testing.assert_array_equal(d.T, gd)
        np.testing.assert_array_equal(ctx.costdists, gcd)
        np.testing.assert_array_equal(ctx.costblocks, gcb)
    finally:
        _lib.check(L.ssg_set_launch(h, C.byref(_lib.LaunchParams(32, 8, -1))), "set launch")


@pytest.fixture(scope="module")
def allknn128(oracle):
    """A 3,000 x 128 graph from the oracle's builder and the oracle's allknn over it."""
    import similaritysearch as ss
    from similaritysearch import datasets
    X = datasets.gmm(3000, DIM, seed=211)
    g = oracle.build_graph("SqL2", X, parallel_block=64, seed=3)
    G = ss.SearchGraph(ss.SqL2, ss.MatrixDatabase.from_rows(X),
                       ss.StaticAdjList.from_csr(g["offs"], g["adj"]), g["hints"],
                       ss.BeamSearch(16, 1.1, 10**6), device=0)
    ai, ad = oracle.allknn_graph("SqL2", X, g["offs"], g["adj"], 8, 16, 1.1, 10**6, nthreads=8)
    return G, ai, ad, X, g


@pytest.mark.parametrize("vbuckets", [0, 1, 32, 512, -1])
def test_visited_table_sizes_bit_exact(oracle, g128, allknn128, vbuckets):
    """Wide beams (cap 256) with a maxvisits cut, and allknn over a graph (seeds skipped
    when already visited, one visited set per object across its searches): every LDS table
    size gives the oracle's bits."""
    import similaritysearch as ss
    from similaritysearch import _lib
    X, Q, G, offs = g128
    L = _lib.load()
    h = G._handle()
    Ga, ai, ad, _, _ = allknn128
    for hh in (h, Ga._handle()):
        _lib.check(L.ssg_set_launch(hh, C.byref(_lib.LaunchParams(32, 8, vbuckets))), "set launch")
    try:
        G.algo = ss.BeamSearch(128, 1.1, 900)
        ctx = ss.SearchGraphContext(beamcap=256)
        ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(Q), 10)
        gi, gd, gcd, gcb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 10, 128,
                                             1.1, 900, beamcap=256, nthreads=8)
        np.testing.assert_array_equal(ids.T, gi)
        np.testing.assert_array_equal(d.T, gd)
        np.testing.assert_array_equal(ctx.costdists, gcd)
        np.testing.assert_array_equal(ctx.costblocks, gcb)
        ids, d = ss.allknn(Ga, ss.SearchGraphContext(), 8)
        np.testing.assert_array_equal(ids.T, ai)
        np.testing.assert_array_equal(d.T, ad)
    finally:
        for hh in (h, Ga._handle()):
            _lib.check(L.ssg_set_launch(hh, C.byref(_lib.LaunchParams(32, 8, -1))), "set launch")


def test_launch_autotuner(oracle, g128):
    """ssg_optimize_launch over a mixed batch (per-query k and bsize) in the BF16 form:
    the shape it installs is one of the grid, the rerank it picks keeps the recall it
    reports, and that recall is the one recomputed from a search with the tuned knobs."""
    import similaritysearch as ss
    from similaritysearch import _lib, datasets
    from similaritysearch.device import DeviceArray, Stream
    X, _, G, offs = g128
    L = _lib.load()
    h = G._handle()
    rng = np.random.default_rng(208)
    Q = datasets.gmm(1200, DIM, seed=209)
    m = len(Q)
    kq = rng.choice([1, 10, 50], m).astype(np.int32)
    bq = rng.choice([16, 32, 64], m).astype(np.int32)
    gold, _ = oracle.exhaustive("SqL2", X, Q, 10, nthreads=8)
    dQ, dk, db = DeviceArray.from_numpy(Q), DeviceArray.from_numpy(kq), DeviceArray.from_numpy(bq)
    qo = ss.QueryOpts(dk.ptr, db.ptr, None, None, 0, None)
    bp = ss.BeamParams(16, 1.1, 10**6, 64, 1, 12, 1)
    lp = _lib.LaunchParams()
    rec, ms = C.c_double(), C.c_float()
    gold = np.ascontiguousarray(gold, np.uint32)
    try:
        _lib.check(L.ssg_optimize_launch(h, dQ.ptr, m, 50, 50, C.byref(bp), C.byref(qo),
                                         gold.ctypes.data, 10, 0.95, C.byref(lp), C.byref(rec),
                                         C.byref(ms)), "optimize launch")
        assert lp.slots_per_cu in (8, 16, 24, 32, 48, 64) and lp.rounds in (4, 8, 10)
        assert lp.vbuckets in (-1, 0)
        assert bp.rerank in (12, 14, 16, 18, 24) and bp.expand in (1, 2) and ms.value > 0
        got = _lib.LaunchParams()
        _lib.check(L.ssg_get_launch(h, C.byref(got)), "get launch")
        assert (got.slots_per_cu, got.rounds, got.vbuckets) == (lp.slots_per_cu, lp.rounds,
                                                                lp.vbuckets)
        dI, dD = DeviceArray((m, 50), np.uint32), DeviceArray((m, 50), np.float32)
        st = Stream()
        _lib.check(L.ssg_beamsearch_dev(h, dQ.ptr, m, 50, 50, C.byref(bp), C.byref(qo), dI.ptr,
                                        dD.ptr, None, None, st.handle), "search")
        st.synchronize()
        ids = dI.numpy()
        kk = np.minimum(kq, 10)
        r = np.mean([len(set(ids[j, :kk[j]]) & set(gold[j, :kk[j]])) / kk[j] for j in range(m)])
        assert r == pytest.approx(rec.value)
    finally:
        _lib.check(L.ssg_set_launch(h, C.byref(_lib.LaunchParams(32, 8, -1))), "set launch")


def test_mlp_gradients_at_learned_shape(oracle):
    """D 128 -> 256 -> 256 -> 4096 + 8, batch 1024: the output layer's forward product
    (1024 x 4104 x 256) takes the 128-tile GEMM.  Loss within 2e-3 relative and every
    gradient within 2e-2 of the layer's largest, against the FP64 oracle that rounds the
    same tensors to FP16."""
    import mlp_oracle as M
    from similaritysearch.learned import Mlp
    rng = np.random.default_rng(205)
    D, H, L, NB, B = 128, 256, 4096, 8, 1024
    mlp = Mlp(D, H, H, L, NB, seed=6)
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
        assert np.abs(gW - rgW[l]).max() <= 2e-2 * np.abs(rgW[l]).max()
        assert np.abs(gb - rgb[l]).max() <= 2e-2 * np.abs(rgb[l]).max() + 1e-7


@pytest.fixture(scope="module")
def learned128(g128):
    import similaritysearch as ss
    from similaritysearch import datasets
    from similaritysearch.learned import LearnedSearch
    X, Q, G, offs = g128
    G.algo = ss.BeamSearch(16, 1.1, 10**6)
    LS = LearnedSearch(G, nlandmarks=4096, hidden=256)
    LS.fit(datasets.gmm(6000, DIM, seed=206), ss.SearchGraphContext(), epochs=4, batch=512)
    return LS


def _np_forward(mlp, Q):
    """FP64 forward with the FP16 roundings of the device path (inputs, weights and hidden
    activations are FP16 operands)."""
    h16 = lambda a: a.astype(np.float16).astype(np.float64)
    Ws = [mlp.get(l) for l in range(3)]
    a = h16(Q)
    for l in range(2):
        a = h16(np.maximum(a @ h16(Ws[l][0]).T + Ws[l][1], 0))
    return a @ h16(Ws[2][0]).T + Ws[2][1]


def test_learned_inference_and_search_dim128(oracle, g128, learned128):
    """Inference over 2,048 queries (the 128-tile GEMM) against a NumPy forward, then the
    planned BF16 search bit-exact against the oracle with the same per-query hints and beam
    sizes, and the fused ssg_learned_search_dev equal to that composition."""
    import similaritysearch as ss
    from similaritysearch import _lib, datasets
    from similaritysearch.device import DeviceArray, Stream
    X, _, G, offs = g128
    LS = learned128
    Q = datasets.gmm(2048, DIM, seed=207)
    m, nh, k, rr = len(Q), 2, 10, 14
    L = _lib.load()
    st = Stream()
    dQ = DeviceArray.from_numpy(Q)
    dH = DeviceArray((m, nh), np.uint32)
    dB = DeviceArray((m,), np.int32)
    dC = DeviceArray((m,), np.float32)
    _lib.check(L.ssg_mlp_infer_dev(LS.mlp._p, dQ.ptr, m, LS.dlandmarks.ptr, nh, 0.0, dH.ptr, dB.ptr,
                                   dC.ptr, st.handle), "infer")
    st.synchronize()
    H, Bk, Cf = dH.numpy(), dB.numpy(), dC.numpy()
    z = _np_forward(LS.mlp, Q)
    zl, zb = z[:, :LS.L], z[:, LS.L:]
    p = np.exp(zl - zl.max(1, keepdims=True))
    conf = p.max(1) / p.sum(1)
    # FP32 accumulation order and FP16 activation roundings: confidences within 2e-3; the
    # top landmark and the bucket agree wherever the runner-up trails by 0.02 or more
    assert np.abs(Cf - conf).max() <= 2e-3
    srt = np.sort(zl, axis=1)
    clear = srt[:, -1] - srt[:, -2] >= 0.02
    top = LS.landmarks[zl.argmax(1)]
    np.testing.assert_array_equal(H[clear, 0], top[clear])
    sb = np.sort(zb, axis=1)
    bclear = sb[:, -1] - sb[:, -2] >= 0.02
    np.testing.assert_array_equal(Bk[bclear], zb.argmax(1)[bclear])
    assert clear.mean() > 0.9 and bclear.mean() > 0.9

    table = LS.table(0)
    dN = DeviceArray((m,), np.int32)
    dS = DeviceArray((m,), np.int32)
    min_conf = float(np.quantile(Cf, 0.1))  # a tenth of the queries take the plain path
    _lib.check(L.ssg_learned_plan_dev(G._handle(), dB.ptr, dC.ptr, m, nh, min_conf,
                                      table.ctypes.data, len(table), 16, dN.ptr, dS.ptr,
                                      st.handle), "plan")
    st.synchronize()
    nhq, bsq = dN.numpy(), dS.numpy()
    assert (nhq == -1).sum() > 0 and (nhq == nh).sum() > 0
    bp = ss.BeamParams(16, 1.1, 10**6, 256, 1, rr, 1)
    qo = ss.QueryOpts(None, dS.ptr, dH.ptr, dN.ptr, nh, None)
    dI, dD = DeviceArray((m, k), np.uint32), DeviceArray((m, k), np.float32)
    dcd, dcb = DeviceArray((m,), np.int64), DeviceArray((m,), np.int64)
    _lib.check(L.ssg_beamsearch_dev(G._handle(), dQ.ptr, m, k, k, C.byref(bp), C.byref(qo), dI.ptr,
                                    dD.ptr, dcd.ptr, dcb.ptr, st.handle), "search")
    st.synchronize()
    oi, od, ocd, ocb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, k, 16, 1.1,
                                         10**6, beamcap=256, bq=bsq, qhints=H, qh_count=nhq,
                                         nthreads=8, bf16_rerank=rr)
    np.testing.assert_array_equal(dI.numpy(), oi)
    np.testing.assert_array_equal(dD.numpy(), od)
    np.testing.assert_array_equal(dcd.numpy(), ocd)
    np.testing.assert_array_equal(dcb.numpy(), ocb)

    fI, fD = DeviceArray((m, k), np.uint32), DeviceArray((m, k), np.float32)
    fcd, fcb = DeviceArray((m,), np.int64), DeviceArray((m,), np.int64)
    _lib.check(L.ssg_learned_search_dev(G._handle(), LS.mlp._p, LS.dlandmarks.ptr, dQ.ptr, m, k, k,
                                        C.byref(bp), nh, min_conf, 0.0, table.ctypes.data,
                                        len(table), fI.ptr, fD.ptr, fcd.ptr, fcb.ptr, st.handle),
               "learned search")
    st.synchronize()
    np.testing.assert_array_equal(fI.numpy(), oi)
    np.testing.assert_array_equal(fD.numpy(), od)
    np.testing.assert_array_equal(fcd.numpy(), ocd)
