"""The batching form: one persistent launch over a heterogeneous batch (per-query k and
beam size, as searchbatch!(index, ctx, Q, knns) with KnnSorted queues of mixed sizes,
src/SimilaritySearch.jl:384-398), the longest-first processing order, and HIP-graph
capture and replay of the launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_heterogeneous_batch_order_and_graph_replay(oracle):
    import similaritysearch as ss
    from similaritysearch import _lib, datasets
    from similaritysearch.device import DeviceArray, Stream
    rng = np.random.default_rng(41)
    X = datasets.gmm(6000, 16, seed=41)
    Q = datasets.gmm(700, 16, seed=42)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=64, seed=2)
    m = len(Q)
    kq = rng.choice([1, 10, 50], m).astype(np.int32)
    bq = rng.choice([16, 32, 64, 128, 256], m).astype(np.int32)
    offs = np.concatenate([[0], G.adj.offset]).astype(np.int64)
    oi, od, ocd, ocb = oracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, 10, 16, 1.1,
                                         10**6, beamcap=256, kq=kq, bq=bq)
    L = _lib.load()
    h = G._handle()
    st = Stream()
    dQ, dk, db = DeviceArray.from_numpy(Q), DeviceArray.from_numpy(kq), DeviceArray.from_numpy(bq)
    dI, dD = DeviceArray((m, 50), np.uint32), DeviceArray((m, 50), np.float32)
    dcd, dcb, dO = DeviceArray((m,), np.int64), DeviceArray((m,), np.int64), DeviceArray((m,), np.uint32)
    bp = ss.BeamParams(16, 1.1, 10**6, 256, 0, 0, 1)

    def run(order):
        qo = ss.QueryOpts(dk.ptr, db.ptr, None, None, 0, dO.ptr if order else None)
        _lib.check(L.ssg_beamsearch_dev(h, dQ.ptr, m, 50, 50, C.byref(bp), C.byref(qo), dI.ptr,
                                        dD.ptr, dcd.ptr, dcb.ptr, st.handle), "searchbatch!")

    def check():
        np.testing.assert_array_equal(dI.numpy(), oi)
        np.testing.assert_array_equal(dD.numpy(), od)
        np.testing.assert_array_equal(dcd.numpy(), ocd)
        np.testing.assert_array_equal(dcb.numpy(), ocb)

    run(False)
    st.synchronize()
    check()
    # longest-first order: a permutation whose log2 classes of bsize never increase
    _lib.check(L.ssg_order_by_bsize_dev(h, db.ptr, m, dO.ptr, st.handle), "order")
    st.synchronize()
    order = dO.numpy()
    np.testing.assert_array_equal(np.sort(order), np.arange(m))
    cls = np.floor(np.log2(bq[order])).astype(int)
    assert np.all(np.diff(cls) <= 0)
    run(True)
    st.synchronize()
    check()
    # recorded once, replayed three times; every replay starts from cleared outputs
    _lib.check(L.ssg_capture_begin(st.handle), "capture")
    try:
        run(True)
    finally:
        g = C.c_void_p()
        rc = L.ssg_capture_end(st.handle, C.byref(g))
    _lib.check(rc, "capture end")
    try:
        for _ in range(3):
            dI.upload(np.zeros((m, 50), np.uint32))
            dD.upload(np.zeros((m, 50), np.float32))
            _lib.check(L.ssg_graph_launch(g, st.handle), "replay")
            st.synchronize()
            check()
    finally:
        L.ssg_graph_destroy(g)
