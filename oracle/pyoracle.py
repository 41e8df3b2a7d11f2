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
# the same restatement built with the reference's FP32 SIMD distance, for timing the CPU
# baseline only (never a checker: its rounding differs from the canonical distance)
_BASE_PATH = os.path.join(_HERE, "_build", "libssg_cpubase.so")
_libs = {}

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


def lib(baseline=False):
    path = _BASE_PATH if baseline else _LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
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
        L.or_beamsearch.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _p, _p, _i64, _p, _p, _i32, _p,
                                    _p, _i64, _i32, _p, _i32, _i32, _p, _f32, _i64, _i32, _p,
                                    _p, _p, _p, C.c_int]
        L.or_beamsearch_bf16.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _p, _p, _i64, _p, _p,
                                         _i32, _p, _p, _i64, _i32, _p, _i32, _i32, _p, _f32,
                                         _i64, _i32, _i32, _p, _p, _p, _p, C.c_int]
        L.or_bf16_round.restype = C.c_float
        L.or_bf16_round.argtypes = [C.c_float]
        L.or_dist_bf16.restype = C.c_float
        L.or_dist_bf16.argtypes = [C.c_int, _p, _p, C.c_int]
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
        L.or_build_graph_staged.restype = C.c_int
        L.or_build_graph_staged.argtypes = [C.c_int, _p, _i64, C.c_int, C.POINTER(BuildParams), _p,
                                            C.c_int, C.POINTER(Graph)]
        L.or_free.argtypes = [_p]
        L.or_hsp_queries.argtypes = [C.c_int, _p, C.c_int, _i64, _i32, _p, _p, _p, _p, C.c_int]
        L.or_fp8_mfma.argtypes = [_p, _i64, _p, _i64, C.c_int, _p, C.c_int, C.c_int, C.c_int, _p, _p]
        L.or_fma_norms.argtypes = [_p, _i64, C.c_int, _p]
        L.or_optimize_minrecall.restype = C.c_double
        L.or_optimize_minrecall.argtypes = [C.c_int, _p, _i64, C.c_int, _p, _p, _p, _i64, _p,
                                            _i64, _i32, C.c_float, _i64, _i32, _p, _p, _p,
                                            C.c_int]
        _libs[path] = L
    return _libs[path]


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


def exhaustive(metric, X, Q, k, nthreads=1, baseline=False):
    """Restated ExhaustiveSearch over every query.  baseline=True runs the CPU-baseline
    build (the reference's FP32 SIMD distance; timing only)."""
    X, Q = _f32c(X), _f32c(Q)
    m = Q.shape[0]
    ids = np.zeros((m, k), np.uint32)
    d = np.full((m, k), np.inf, np.float32)
    lib(baseline).or_exhaustive(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1], _ptr(Q), m,
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
               bq=None, qhints=None, qh_count=None, nthreads=1, qvisit=None, bf16_rerank=None,
               baseline=False):
    """Restated searchbatch over a SearchGraph.  Returns ids, dists, costdists, costblocks.
    qvisit: per query a 1-based object marked visited before its search (0 = none).
    bf16_rerank (an int): the BF16 form instead (no reference counterpart): the search runs
    on dist_bf16 with a result queue of max(k, bf16_rerank) entries, reranked with the
    canonical distance before the first k are written.  baseline=True runs the CPU-baseline
    build (the reference's FP32 SIMD distance; timing only)."""
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
    qv = None if qvisit is None else np.ascontiguousarray(qvisit, np.uint32)
    ids = np.zeros((m, kstride), np.uint32)
    d = np.full((m, kstride), np.inf, np.float32)
    cd = np.zeros(m, np.int64)
    cb = np.zeros(m, np.int64)
    common = (METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1], _ptr(offs),
              _ptr(adj), _ptr(hints), len(hints), _ptr(qh), _ptr(qc),
              0 if qh is None else qh.shape[1], _ptr(qv), _ptr(Q), m, int(k), _ptr(kq_a), kstride,
              int(bsize), _ptr(bq_a), float(delta), int(maxvisits), int(beamcap))
    if bf16_rerank is None:
        lib(baseline).or_beamsearch(*common, _ptr(ids), _ptr(d), _ptr(cd), _ptr(cb), nthreads)
    else:
        lib().or_beamsearch_bf16(*common, int(bf16_rerank), _ptr(ids), _ptr(d), _ptr(cd),
                                 _ptr(cb), nthreads)
    return ids, d, cd, cb


def unit_rows(R):
    """Cosine (src/dist/cos.jl:103) as searched: each row divided by its FP64 Euclidean
    norm and rounded to Float32 once (zero rows stay zero); the canonical NormCosine
    distance of such rows is the cosine distance.  Pinned in tests/test_oracle.py by the
    closed forms of test/testdistances.jl:61-76 and an FP64 bound."""
    R = np.asarray(R, np.float64)
    out = np.zeros(R.shape, np.float32)
    for i, r in enumerate(R):
        nr = np.sqrt(np.dot(r, r))
        if nr > 0:
            out[i] = (r / nr).astype(np.float32)
    return out


def dist_bf16(metric, x, q):
    """The BF16 form's search distance: row x rounded to BF16, FP32 query and arithmetic
    in the canonical run/tree order (oracle/ssg_oracle.c, or_dist_bf16)."""
    x, q = _f32c(x), _f32c(q)
    return lib().or_dist_bf16(METRICS.get(metric, metric), _ptr(x), _ptr(q), x.shape[-1])


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
                opt_ksearch=10, beamcap=32, seed=0, stages=None):
    """index! restated; returns dict(offs, adj, hints, bsize, delta, maxvisits).  stages
    (increasing object counts ending at n): index! over the first stages[0] objects, then
    append_items! up to each later count (the graph built by successive calls)."""
    X = _f32c(X)
    p = BuildParams(nb_logbase, nb_minsize, parallel_block, hints_logbase, cb_logbase,
                    cb_starting, opt_minrecall, opt_numqueries, opt_ksearch, beamcap, seed)
    g = Graph()
    if stages is None:
        rc = lib().or_build_graph(METRICS.get(metric, metric), _ptr(X), X.shape[0], X.shape[1],
                                  C.byref(p), C.byref(g))
    else:
        ends = np.ascontiguousarray(stages, np.int64)
        rc = lib().or_build_graph_staged(METRICS.get(metric, metric), _ptr(X), X.shape[0],
                                         X.shape[1], C.byref(p), _ptr(ends), len(ends), C.byref(g))
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


def fp8_scale_exp(a):
    """The FP8 operand scale of a set of values: the largest integer e with
    max |a| * 2^e <= 448 (the largest finite e4m3fn value), clamped to [-120, 120]; 0 when
    every value is 0 or one is not finite.  Pinned by hand cases in tests/test_oracle.py."""
    a = np.asarray(a, np.float32)
    mx = float(np.max(np.abs(a))) if a.size else 0.0
    if mx == 0.0 or not np.isfinite(mx):
        return 0
    f, E = np.frexp(mx)          # mx = f * 2^E, f in [0.5, 1); 448 = 0.875 * 2^9
    e = (9 if f <= 0.875 else 8) - int(E)
    return min(120, max(-120, e))


def fp8_round(x, e):
    """float32 values -> the OCP FP8 e4m3fn value nearest x * 2^e (round to nearest, ties
    to even; 3 mantissa bits, exponent bias 7, subnormal quantum 2^-9), times 2^-e, as
    float64.  Needs |x * 2^e| <= 448.  Pinned against torch's float8_e4m3fn conversion and
    hand cases in tests/test_oracle.py."""
    y = np.ldexp(np.asarray(x, np.float32).astype(np.float64), e)
    a = np.abs(y)
    if (a > 448.0).any():
        raise ValueError("value beyond the e4m3fn range")
    _, E = np.frexp(np.where(a > 0, a, 1.0))   # a in [2^(E-1), 2^E)
    q = np.ldexp(1.0, np.maximum(E - 1, -6) - 3)  # spacing of the binade (subnormals: 2^-9)
    return np.ldexp(np.round(y / q) * q, -e)     # np.round: ties to even


def _prec_operands(X, Q, fmt):
    if fmt == "bf16":
        return bf16_round(X).astype(np.float64), bf16_round(Q).astype(np.float64)
    if fmt == "fp8":
        return fp8_round(X, fp8_scale_exp(X)), fp8_round(Q, fp8_scale_exp(Q))
    raise ValueError(fmt)


def prec_approx(metric, X, Q, fmt="bf16"):
    """(m, n) FP64 approximate distances of the candidate pass: s = sum of the exact
    products of the rounded coordinates (fmt "bf16": BF16; "fp8": e4m3fn of X * 2^ex and
    Q * 2^eq, ex and eq from fp8_scale_exp of X and of Q, scaled back); NormCosine 1 - s,
    SqL2 |q|^2 + |x|^2 - 2 s (norms of the FP32 vectors), L2 its square root."""
    mt = METRICS.get(metric, metric)
    Xb, Qb = _prec_operands(X, Q, fmt)
    s = Qb @ Xb.T
    if mt == NORMCOS:
        return 1.0 - s
    xn = (np.asarray(X, np.float64) ** 2).sum(1)
    qn = (np.asarray(Q, np.float64) ** 2).sum(1)
    d = qn[:, None] + xn[None, :] - 2.0 * s
    return np.sqrt(np.maximum(d, 0.0)) if mt == L2 else d


def prec_error_bound(metric, X, Q):
    """(m, n) bound on |GPU approximate distance - prec_approx| of the BF16 pass for the
    SqL2/NormCosine forms: FP32 accumulation of the dim exact BF16 products plus the
    -|x|^2/2 start value in any order ((dim + 2) units of 2^-24 of the absolute sum),
    FP32 norms of dim terms, and the final FP32 roundings.  (The FP8 pass does not
    accumulate in FP32 order: fp8_scores restates it exactly.)"""
    mt = METRICS.get(metric, metric)
    u = 2.0 ** -24
    dim = np.asarray(X).shape[1]
    Xb, Qb = _prec_operands(X, Q, "bf16")
    Xb, Qb = np.abs(Xb), np.abs(Qb)
    a = Qb @ Xb.T
    if mt == NORMCOS:
        return (dim + 2) * u * a + 2 * u
    xn = (np.asarray(X, np.float64) ** 2).sum(1)
    qn = (np.asarray(Q, np.float64) ** 2).sum(1)
    acc = (dim + 2) * u * (2.0 * a + xn[None, :])
    norms = (dim + 1) * u * (xn[None, :] + qn[:, None])
    return acc + norms + 4 * u * (qn[:, None] + xn[None, :] + 2.0 * a)


def fma_norms(X):
    """FP32 squared norms of the rows as the device computes them (fma chain in
    increasing index)."""
    X = _f32c(X)
    out = np.zeros(len(X), np.float32)
    lib().or_fma_norms(_ptr(X), len(X), X.shape[1], _ptr(out))
    return out


def fp8_operands(X, e, K=128):
    """Rows of X as the FP8 pass holds them: the e4m3fn values of X * 2^e (fp8_round, in
    scaled units, exact in float32), zero padded to K values."""
    X = np.asarray(X, np.float32)
    out = np.zeros((len(X), K), np.float32)
    out[:, :X.shape[1]] = np.ldexp(fp8_round(X, e), e)
    return out


FP8_ACC_UNITS = 18  # accumulation bound of the FP8 product, units of 2^-24 of the magnitude


def fp8_mfma(P, R, C=None, c_mode=2, s=0, variant=0):
    """The e4m3fn matrix-core product of P rows and R rows (exact FP8 values, K a
    multiple of 8 each) started at C (c_mode 0: (m, n); 1: per R row; 2: zero), operand
    scale 2^s: returns (value, magnitude), both (m, n) float64, with the device's FP32
    result within FP8_ACC_UNITS * 2^-24 * magnitude of value; see or_fp8_mfma."""
    P, R = _f32c(P), _f32c(R)
    m, n, K = len(P), len(R), P.shape[1]
    Cc = None if C is None else _f32c(C)
    D = np.zeros((m, n), np.float64)
    M = np.zeros((m, n), np.float64)
    lib().or_fp8_mfma(_ptr(P), m, _ptr(R), n, K, _ptr(Cc), c_mode, s, variant, _ptr(D), _ptr(M))
    return D, M


def fp8_scores(metric, X, Q, Qscale=None):
    """Scores of the FP8 candidate pass as (value, magnitude), (m, n) float64 (see
    fp8_mfma): the e4m3fn product of the rows scaled by 2^ex (ex = fp8_scale_exp(X)) and
    the queries scaled by 2^eq (eq = fp8_scale_exp of Qscale, every query of the call, or
    of Q), operand scale 2^-(ex + eq), started at -|x|^2/2 (fma_norms) for SqL2/L2 and at
    0 for NormCosine; larger is nearer."""
    mt = METRICS.get(metric, metric)
    ex = fp8_scale_exp(X)
    eq = fp8_scale_exp(Q if Qscale is None else Qscale)
    Px, Pq = fp8_operands(X, ex), fp8_operands(Q, eq)
    if mt == NORMCOS:
        return fp8_mfma(Pq, Px, None, 2, -(ex + eq))
    c = (np.float32(-0.5) * fma_norms(X)).astype(np.float32)
    return fp8_mfma(Pq, Px, c, 1, -(ex + eq))


def fp8_approx(metric, X, Q, Qscale=None):
    """(m, n) float64 approximate distances of the FP8 pass and their bound: NormCosine
    1 - s, SqL2 |q|^2 - 2 s (fma_norms of the query; the device evaluates both in FP32
    from its FP32 score s), with |device - value| <= bound (the score's accumulation bound
    and the final FP32 rounding)."""
    mt = METRICS.get(metric, metric)
    if mt == L2:
        raise ValueError("the FP8 approximate-distance bound covers SqL2 and NormCosine")
    S, M = fp8_scores(metric, X, Q, Qscale)
    u = 2.0 ** -24
    eb = FP8_ACC_UNITS * u * M
    if mt == NORMCOS:
        return 1.0 - S, eb + u * (1.0 + np.abs(S) + eb)
    qn = fma_norms(Q).astype(np.float64)[:, None]
    return qn - 2.0 * S, 2.0 * eb + u * (qn + 2.0 * (np.abs(S) + eb))


# ------------------------------------------------------------------ batch consumers
def bichromatic_closestpair(metric, A, B, min_k=8, samedata=False):
    """The closest pair (i of A, j of B, dist) as bichromatic_closestpair over an
    ExhaustiveSearch of A (src/bichromatic/bichromaticclosestpair.jl:52-139): each object
    of B takes its min_k nearest in (distance, id) order; with samedata a nearest entry
    equal to the query is dropped (pop_min!); the least distance wins, ties to the lowest
    j (strict < in object order, first batch on equal minima).  Pinned by brute force
    over all pairs in tests/test_oracle.py."""
    ids, d = exhaustive(metric, A, B, min_k)
    return closest_of_rows(ids, d, samedata)


def closest_of_rows(ids, d, samedata):
    """The pair bichromatic_closestpair picks from each object's (m, k) result row
    (bichromaticclosestpair.jl:75-81, 103-106): drop the row's nearest entry when it is
    the object itself (samedata), then the least distance, lowest j on ties."""
    best = (0, 0, np.float32(np.inf))
    k = ids.shape[1]
    for j in range(ids.shape[0]):
        t = 1 if samedata and ids[j, 0] == j + 1 else 0
        if t < k and ids[j, t] != 0 and d[j, t] < best[2]:
            best = (int(ids[j, t]), j + 1, d[j, t])
    return best


def neardup(metric, X, eps, k=8, blocksize=256, filterblocks=True, nthreads=1, graph=None):
    """neardup over an exact index (src/neardup.jl:69-256): blocks of blocksize objects;
    a block's objects search the net built so far (nearest in (distance, id) order; a
    near duplicate when within eps), the rest are filtered in order against the block's
    new centers (first center at the least distance), which then join the net.  Returns
    (map, nn, dist, centers, costdists) as in similaritysearch.neardup.  Pinned in
    tests/test_oracle.py: one block is the greedy eps-net in object order (brute force),
    and blocked runs keep the net invariants.  graph (a dict of build_graph arguments):
    the recall < 1 form (src/neardup.jl:71-76), the net a SearchGraph built by index! and
    grown by append_items! of each block's new centers (build_graph with stages), searched
    by the restated beam search with KnnSorted(k); costdists counts those searches'
    evaluations instead of the net's length per object."""
    X = _f32c(X)
    eps = np.float32(eps)
    n = len(X)
    blocksize = min(blocksize, n)
    L = np.zeros(n, np.int32)
    D = np.zeros(n, np.float32)
    M, net = [], []
    cost = 0

    def block(items):
        nonlocal cost
        if not filterblocks:
            for i in items:
                M.append(i); L[i - 1] = i; D[i - 1] = 0.0
            return list(items)
        P = dist_batch(metric, X[np.asarray(items) - 1], X[np.asarray(items) - 1])
        tmp = [0]
        M.append(items[0]); L[items[0] - 1] = items[0]; D[items[0] - 1] = 0.0
        for ii in range(1, len(items)):
            cost += len(tmp)
            best_t, best_d = -1, np.float32(np.inf)
            for t in tmp:
                if P[ii, t] < best_d:
                    best_t, best_d = t, P[ii, t]
            i = items[ii]
            if best_d > eps:
                tmp.append(ii); M.append(i); L[i - 1] = i; D[i - 1] = 0.0
            else:
                L[i - 1] = items[best_t]; D[i - 1] = best_d
        return [items[t] for t in tmp]

    ends = []  # net sizes after each append_items! (graph form)
    for r0 in range(0, n, blocksize):
        rng = list(range(r0 + 1, min(n, r0 + blocksize) + 1))
        if not net:
            new = block(rng)
        elif graph is not None:
            Xn = X[np.asarray(net) - 1]
            g = build_graph(metric, Xn, stages=ends, **graph)
            ids, d, cd, _ = beamsearch(metric, Xn, g["offs"], g["adj"], g["hints"],
                                       X[r0:r0 + len(rng)], k, g["bsize"], g["delta"],
                                       g["maxvisits"], beamcap=graph.get("beamcap", 32),
                                       nthreads=nthreads)
            cost += int(cd.sum())
            imap = []
            for t, j in enumerate(rng):
                if ids[t, 0] == 0 or d[t, 0] > eps:
                    imap.append(j)
                else:
                    D[j - 1] = d[t, 0]; L[j - 1] = M[int(ids[t, 0]) - 1]
            new = block(imap) if imap else []
        else:
            # searchbatch! of the block against the net (src/neardup.jl:121-127); each exact
            # search charges the net's length (src/exact/sequential-exhaustive.jl:78)
            cost += len(rng) * len(net)
            ids, d = exhaustive(metric, X[np.asarray(net) - 1], X[r0:r0 + len(rng)], 1, nthreads)
            imap = []
            for t, j in enumerate(rng):
                if ids[t, 0] == 0 or d[t, 0] > eps:
                    imap.append(j)
                else:
                    D[j - 1] = d[t, 0]; L[j - 1] = M[int(ids[t, 0]) - 1]
            new = block(imap) if imap else []
        net.extend(new)
        if new:
            ends.append(len(net))
    return np.asarray(M, np.uint32), L, D, np.asarray(net, np.uint32), cost


def hsp_queries(metric, X, Q, ids, dists, nthreads=1, baseline=False):
    """hsp_queries (src/hsp.jl:8-33, 77-101) over (m, k) candidate rows (0 ends a row): a
    candidate is kept unless an already kept one is strictly nearer to it than its given
    distance (canonical distances); the kept enter a KnnSorted of capacity k.  Returns
    (m, k) kept ids and distances, then 0 / Inf (or_hsp_queries in ssg_oracle.c).  Pinned
    by the keep/drop invariants in tests/test_oracle.py.  baseline=True runs the CPU-baseline
    build (the reference's FP32 SIMD distance; timing only)."""
    X = _f32c(X)
    ids = np.ascontiguousarray(ids, np.uint32)
    dd = np.ascontiguousarray(dists, np.float32)
    m, k = ids.shape
    oi = np.zeros((m, k), np.uint32)
    od = np.full((m, k), np.inf, np.float32)
    lib(baseline).or_hsp_queries(METRICS.get(metric, metric), _ptr(X), X.shape[1], m, k, _ptr(ids), _ptr(dd),
                         _ptr(oi), _ptr(od), nthreads)
    return oi, od
