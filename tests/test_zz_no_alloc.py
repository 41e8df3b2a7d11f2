"""ssg_beamsearch_noalloc_dev (the form of ssg_beamsearch_dev that allocates nothing, its
scratch in a static device array) against ssg_beamsearch_dev: bit-exact ids, distances and
cost counters, on two calls in a row with the second one smaller, in the FP32 and BF16
forms and with per-query k, bsize and processing order."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def index():
    import similaritysearch as ss
    from similaritysearch import datasets
    X = datasets.gmm(12000, 64, seed=301)
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), parallel_block=128, seed=3)
    return G


def _run(G, fn, Q, kq, bq, prec, stream):
    import similaritysearch as ss
    from similaritysearch import _lib
    from similaritysearch.device import DeviceArray
    L = _lib.load()
    m = len(Q)
    dQ, dk, db = DeviceArray.from_numpy(Q), DeviceArray.from_numpy(kq), DeviceArray.from_numpy(bq)
    dO = DeviceArray((m,), np.uint32)
    _lib.check(L.ssg_order_by_bsize_dev(G._handle(), db.ptr, m, dO.ptr, stream.handle), "order")
    dI, dD = DeviceArray((m, 50), np.uint32), DeviceArray((m, 50), np.float32)
    dcd, dcb = DeviceArray((m,), np.int64), DeviceArray((m,), np.int64)
    bp = ss.BeamParams(16, 1.1, 10**6, 256, prec, 20 if prec else 0, 1)
    qo = ss.QueryOpts(dk.ptr, db.ptr, None, None, 0, dO.ptr)
    _lib.check(getattr(L, fn)(G._handle(), dQ.ptr, m, 50, 50, C.byref(bp), C.byref(qo), dI.ptr,
                              dD.ptr, dcd.ptr, dcb.ptr, stream.handle), fn)
    stream.synchronize()
    return dI.numpy(), dD.numpy(), dcd.numpy(), dcb.numpy()


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_equals_beamsearch_dev_twice(index, prec):
    from similaritysearch import datasets
    from similaritysearch.device import Stream
    G = index
    st = Stream()
    rng = np.random.default_rng(302 + prec)
    for m in (1500, 333):  # the second call is smaller and reuses the scratch
        Q = datasets.gmm(m, 64, seed=int(rng.integers(1 << 30)))
        kq = rng.choice([1, 10, 50], m).astype(np.int32)
        bq = rng.choice([16, 32, 64, 128, 256], m).astype(np.int32)
        ref = _run(G, "ssg_beamsearch_dev", Q, kq, bq, prec, st)
        got = _run(G, "ssg_beamsearch_noalloc_dev", Q, kq, bq, prec, st)
        for r, g in zip(ref, got):
            np.testing.assert_array_equal(g, r)


def test_noalloc_argument_errors(index):
    import similaritysearch as ss
    from similaritysearch import _lib
    from similaritysearch.device import DeviceArray
    L = _lib.load()
    dQ = DeviceArray.from_numpy(np.zeros((4, 64), np.float32))
    dI, dD = DeviceArray((4, 300), np.uint32), DeviceArray((4, 300), np.float32)
    bp = ss.BeamParams(16, 1.1, 10**6, 32, 0, 0, 1)
    rc = L.ssg_beamsearch_noalloc_dev(index._handle(), dQ.ptr, 4, 300, 300, C.byref(bp), None, dI.ptr,
                                      dD.ptr, None, None, None)
    assert rc == -5  # SSG_E_LIMIT: k beyond the register queues
    rc = L.ssg_beamsearch_noalloc_dev(index._handle(), None, 4, 10, 10, C.byref(bp), None, dI.ptr,
                                      dD.ptr, None, None, None)
    assert rc == -1
