"""ctypes front-end of the CPU oracle (oracle/ssg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker.  The product package
never imports it.

Arrays follow the reference's layout: a database or query set is a numpy
(n, dim) float32 C-contiguous array (= Julia's column-major (dim, n)), results
are (m, k) arrays whose row j is query j's column of the reference's (k, m)
matrices, ids are 1-based uint32 with 0 for empty slots.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("SSG_ORACLE_LIB", os.path.join(_HERE, "_build", "libssg_oracle.so"))
_lib = None

SQL2, L2, NORMCOS = 0, 1, 2
METRICS = {"sqL2": SQL2, "SqL2": SQL2, "L2": L2, "NormCosine": NORMCOS, "normcos": NORMCOS}

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_f32 = C.c_float


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class BuildParams(C.Structure):
    _fields_ = [("nb_logbase", C.c_float), ("nb_minsize", C.c_int32),
                ("parallel_block", C.c_int32), ("hints_logbase", C.c_float),
                ("cb_logbase", C.c_float), ("cb_starting", C.c_int32),
                ("opt_minrecall", C.c_float), ("opt_numqueries", C.c_int32),
                ("opt_ksearch", C.c_int32), ("beamcap", C.c_int32), ("seed", C.c_uint64)]


class Graph(C.Structure):
    _fields_ = [("offs", C.POINTER(C.c_int64)), ("adj", C.POINTER(C.c_uint32)),
                ("hints", C.POINTER(C.c_uint32)), ("nnz", C.c_int64), ("nhints", C.c_int64),
                ("bsize", C.c_int32), ("delta", C.c_float), ("maxvisits", C.c_int64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_dist.restype = C.c_float
        L.or_dist.argtypes = [C.c_int, _p, _p, C.c_int]
        L.or_dist_f64.restype = C.c_double
        L.or_dist_f64.argtypes = [C.c_int, _p, _p, C.c_int]
        L.or_dist_batch.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _i64, _p]
        L.or_knn_run.restype = C.c_int
        L.or_knn_run.argtypes = [_i32, _i64, _p, _p, _p, _p, _p, _p, _p]
        L.or_exhaustive.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _i64, _i32, _p, _p, C.c_int]
        L.or_parallel_exhaustive.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _i64, _i32, _i32,
                                             _p, _p]
        L.or_beamsearch.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _p, _p, _i64, _p, _p, _i32,
                                    _p, _i64, _i32, _p, _i32, _i32, _p, _f32, _i64, _i32, _p,
                                    _p, _p, _p, C.c_int]
        L.or_allknn_graph.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _p, _i32, _i32, _f32, _i64,
                                      _i32, _p, _p, C.c_int]
        L.or_rerank.argtypes = [C.c_int, _p, C.c_int, _p, _i64, _i32, _p, _p]
        L.or_macrorecall.restype = C.c_double
        L.or_macrorecall.argtypes = [_p, _i32, _p, _i32, _i64, _i32]
        L.or_fft.argtypes = [C.c_int, _p, _i64, C.c_int, _i32, _i64, C.c_uint64, _p, _p, _p, _p]
        L.or_random_hints.restype = C.c_int64
        L.or_random_hints.argtypes = [_i64, _p, _p, C.c_float, C.c_uint64, _p, _i64]
        L.or_build_graph.restype = C.c_int
        L.or_build_graph.argtypes = [C.c_int, _p, _i64, C.c_int, C.POINTER(BuildParams),
                                     C.POINTER(Graph)]
        L.or_free.argtypes = [_p]
        L.or_optimize_minrecall.restype = C.c_double
        L.or_optimize_minrecall.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _p, _p, _i64, _p,
                                            _i64, _i32, C.c_float, _i64, _i32, _p, _p, _p,
                                            C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def dist(metric, a, b):
    a, b = _f32c(a), _f32c(b)
    return lib().or_dist(METRICS.get(metric, metric), _ptr(a), _ptr(b), a.shape[-1])


def dist_f64(metric, a, b):
    a, b = _f32c(a), _f32c(b)
    return lib().or_dist_f64(METRICS.get(metric, metric), _ptr(a), _ptr(b), a.shape[-1])


def dist_batch(metric, X, Q):
    X, Q = _f32c(X), _f32c(Q)
    out = np.empty((Q.shape[0], X.shape[0]), np.float32)
    lib().or_dist_batch(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1], _ptr(Q),
                        Q.shape[0], _ptr(out))
    return out


def knn_run(cap, ops, ids, dists):
    ops = np.ascontiguousarray(ops, np.int8)
    ids = np.ascontiguousarray(ids, np.uint32)
    dists = np.ascontiguousarray(dists, np.float32)
    acc = np.zeros(len(ops), np.uint8)
    oi = np.zeros(cap + 1, np.uint32)
    od = np.zeros(cap + 1, np.float32)
    ln = np.zeros(1, np.int32)
    rc = lib().or_knn_run(cap, len(ops), _ptr(ops), _ptr(ids), _ptr(dists), _ptr(acc), _ptr(oi),
                          _ptr(od), _ptr(ln))
    if rc != 0:
        raise ValueError("pop from an empty queue")
    return oi[:ln[0]].copy(), od[:ln[0]].copy(), acc.astype(bool)


def exhaustive(metric, X, Q, k, nthreads=1):
    X, Q = _f32c(X), _f32c(Q)
    m = Q.shape[0]
    ids = np.zeros((m, k), np.uint32)
    d = np.full((m, k), np.inf, np.float32)
    lib().or_exhaustive(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1], _ptr(Q), m,
                        k, _ptr(ids), _ptr(d), nthreads)
    return ids, d


def parallel_exhaustive(metric, X, Q, k, nbatches):
    X, Q = _f32c(X), _f32c(Q)
    m = Q.shape[0]
    ids = np.zeros((m, k), np.uint32)
    d = np.full((m, k), np.inf, np.float32)
    lib().or_parallel_exhaustive(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1],
                                 _ptr(Q), m, k, nbatches, _ptr(ids), _ptr(d))
    return ids, d


def beamsearch(metric, X, offs, adj, hints, Q, k, bsize, delta, maxvisits, beamcap=32, kq=None,
               bq=None, qhints=None, qh_count=None, nthreads=1):
    """Restated searchbatch over a SearchGraph.  Returns ids, dists, costdists, costblocks."""
    X, Q = _f32c(X), _f32c(Q)
    offs = np.ascontiguousarray(offs, np.int64)
    adj = np.ascontiguousarray(adj, np.uint32)
    hints = np.ascontiguousarray(hints if hints is not None else [], np.uint32)
    m = Q.shape[0]
    kstride = int(k if kq is None else max(int(np.max(kq)), 1))
    kq_a = None if kq is None else np.ascontiguousarray(kq, np.int32)
    bq_a = None if bq is None else np.ascontiguousarray(bq, np.int32)
    qh = None if qhints is None else np.ascontiguousarray(qhints, np.uint32)
    qc = None if qh_count is None else np.ascontiguousarray(qh_count, np.int32)
    ids = np.zeros((m, kstride), np.uint32)
    d = np.full((m, kstride), np.inf, np.float32)
    cd = np.zeros(m, np.int64)
    cb = np.zeros(m, np.int64)
    lib().or_beamsearch(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1], _ptr(offs),
                        _ptr(adj), _ptr(hints), len(hints), _ptr(qh), _ptr(qc),
                        0 if qh is None else qh.shape[1], _ptr(Q), m, int(k), _ptr(kq_a), kstride,
                        int(bsize), _ptr(bq_a), float(delta), int(maxvisits), int(beamcap),
                        _ptr(ids), _ptr(d), _ptr(cd), _ptr(cb), nthreads)
    return ids, d, cd, cb


def allknn_graph(metric, X, offs, adj, k, bsize, delta, maxvisits, beamcap=32, nthreads=1):
    X = _f32c(X)
    offs = np.ascontiguousarray(offs, np.int64)
    adj = np.ascontiguousarray(adj, np.uint32)
    n = X.shape[0]
    ids = np.zeros((n, k), np.uint32)
    d = np.full((n, k), np.inf, np.float32)
    lib().or_allknn_graph(METRICS.get(metric, metric), _ptr(X), n, X.shape[1], _ptr(offs),
                          _ptr(adj), k, int(bsize), float(delta), int(maxvisits), int(beamcap),
                          _ptr(ids), _ptr(d), nthreads)
    return ids, d


def rerank(metric, X, Q, ids, dists):
    X, Q = _f32c(X), _f32c(Q)
    ids = np.array(ids, np.uint32, copy=True, order="C")
    dists = np.array(dists, np.float32, copy=True, order="C")
    lib().or_rerank(METRICS.get(metric, metric), _ptr(X), X.shape[1], _ptr(Q), Q.shape[0],
                    ids.shape[1], _ptr(ids), _ptr(dists))
    return ids, dists


def macrorecall(gold, res, k=None):
    gold = np.ascontiguousarray(gold, np.uint32)
    res = np.ascontiguousarray(res, np.uint32)
    k = gold.shape[1] if k is None else k
    return lib().or_macrorecall(_ptr(gold), gold.shape[1], _ptr(res), res.shape[1],
                                gold.shape[0], k)


def fft(metric, X, k, start=0, seed=0):
    X = _f32c(X)
    n = X.shape[0]
    centers = np.zeros(k, np.uint32)
    nn = np.zeros(n, np.uint32)
    nnd = np.zeros(n, np.float32)
    eps = np.zeros(k, np.float32)
    lib().or_fft(METRICS.get(metric, metric), _ptr(X), n, X.shape[1], k, start, seed,
                 _ptr(centers), _ptr(nn), _ptr(nnd), _ptr(eps))
    return centers, nn, nnd, eps


def random_hints(offs, adj, logbase=1.1, seed=0):
    offs = np.ascontiguousarray(offs, np.int64)
    adj = np.ascontiguousarray(adj, np.uint32)
    n = len(offs) - 1
    cap = int(np.ceil(np.log(max(n, 2)) / np.log(logbase))) + 2
    out = np.zeros(cap, np.uint32)
    nh = lib().or_random_hints(n, _ptr(offs), _ptr(adj), logbase, seed, _ptr(out), cap)
    return out[:nh].copy()


def build_graph(metric, X, parallel_block=1, nb_logbase=1.3, nb_minsize=2, hints_logbase=1.1,
                cb_logbase=1.5, cb_starting=256, opt_minrecall=0.97, opt_numqueries=32,
                opt_ksearch=10, beamcap=32, seed=0):
    """index! restated; returns dict(offs, adj, hints, bsize, delta, maxvisits)."""
    X = _f32c(X)
    p = BuildParams(nb_logbase, nb_minsize, parallel_block, hints_logbase, cb_logbase,
                    cb_starting, opt_minrecall, opt_numqueries, opt_ksearch, beamcap, seed)
    g = Graph()
    rc = lib().or_build_graph(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1],
                              C.byref(p), C.byref(g))
    if rc != 0:
        raise ValueError("empty database")
    n = X.shape[0]
    offs = np.ctypeslib.as_array(g.offs, shape=(n + 1,)).copy()
    adj = np.ctypeslib.as_array(g.adj, shape=(max(g.nnz, 1),))[:g.nnz].copy()
    hints = (np.ctypeslib.as_array(g.hints, shape=(max(g.nhints, 1),))[:g.nhints].copy()
             if g.nhints > 0 else np.zeros(0, np.uint32))
    out = dict(offs=offs, adj=adj, hints=hints, bsize=g.bsize, delta=g.delta,
               maxvisits=g.maxvisits)
    lib().or_free(C.cast(g.offs, C.c_void_p))
    lib().or_free(C.cast(g.adj, C.c_void_p))
    lib().or_free(C.cast(g.hints, C.c_void_p))
    return out


def optimize_minrecall(metric, X, offs, adj, hints, Q, ksearch, minrecall, cur_maxvisits=10**6,
                       beamcap=32, nthreads=1):
    X, Q = _f32c(X), _f32c(Q)
    offs = np.ascontiguousarray(offs, np.int64)
    adj = np.ascontiguousarray(adj, np.uint32)
    hints = np.ascontiguousarray(hints, np.uint32)
    bs = C.c_int32()
    dl = C.c_float()
    mv = C.c_int64()
    r = lib().or_optimize_minrecall(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1],
                                    _ptr(offs), _ptr(adj), _ptr(hints), len(hints), _ptr(Q),
                                    Q.shape[0], ksearch, minrecall, cur_maxvisits, beamcap,
                                    C.byref(bs), C.byref(dl), C.byref(mv), nthreads)
    return dict(bsize=bs.value, delta=dl.value, maxvisits=mv.value, recall=r)


# --------------------------------------------------------------------------- precision form
# The precision form of exhaustive k-NN (a candidate pass on BF16 operands, then rerank!
# with the canonical distance; src/rerank.jl:22-98 for the second step) has no reference
# counterpart for its first step.  Its approximate distance is modelled here exactly in
# FP64 from the BF16-rounded operands; the GPU accumulates the same products in FP32,
# which the tests bound with prec_error_bound.

def bf16_round(x):
    """Float32 -> nearest BF16 value (round to nearest, ties to even), returned as float32.
    Pinned against torch's bfloat16 conversion and hand cases in tests/test_oracle.py."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def prec_approx(metric, X, Q):
    """(m, n) FP64 approximate distances of the BF16 candidate pass: s = sum of the exact
    products of the BF16-rounded coordinates; NormCosine 1 - s, SqL2 |q|^2 + |x|^2 - 2 s
    (norms of the FP32 vectors), L2 its square root."""
    mt = METRICS.get(metric, metric)
    Xb = bf16_round(X).astype(np.float64)
    Qb = bf16_round(Q).astype(np.float64)
    s = Qb @ Xb.T
    if mt == NORMCOS:
        return 1.0 - s
    xn = (np.asarray(X, np.float64) ** 2).sum(1)
    qn = (np.asarray(Q, np.float64) ** 2).sum(1)
    d = qn[:, None] + xn[None, :] - 2.0 * s
    return np.sqrt(np.maximum(d, 0.0)) if mt == L2 else d


def prec_error_bound(metric, X, Q):
    """(m, n) bound on |GPU approximate distance - prec_approx| for the SqL2/NormCosine
    forms: FP32 accumulation of the dim exact BF16 products plus the -|x|^2/2 start value
    in any order ((dim + 2) units of 2^-24 of the absolute sum), FP32 norms of dim terms,
    and the final FP32 roundings."""
    mt = METRICS.get(metric, metric)
    u = 2.0 ** -24
    dim = np.asarray(X).shape[1]
    Xb = np.abs(bf16_round(X).astype(np.float64))
    Qb = np.abs(bf16_round(Q).astype(np.float64))
    a = Qb @ Xb.T
    if mt == NORMCOS:
        return (dim + 2) * u * a + 2 * u
    xn = (np.asarray(X, np.float64) ** 2).sum(1)
    qn = (np.asarray(Q, np.float64) ** 2).sum(1)
    acc = (dim + 2) * u * (2.0 * a + xn[None, :])
    norms = (dim + 1) * u * (xn[None, :] + qn[:, None])
    return acc + norms + 4 * u * (qn[:, None] + xn[None, :] + 2.0 * a)
