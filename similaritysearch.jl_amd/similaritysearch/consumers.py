"""Batch consumers of the search path, on the GPU indexes: closest pairs, near duplicates
and half-space proximal neighborhoods.

Each follows the reference's algorithm step for step; every distance comes from the GPU
(searchbatch!, allknn! or rerank! through the C ABI) with the canonical FP32 distance,
and only the reference's sequential bookkeeping runs on the host.
- closestpair / bichromatic_closestpair: src/bichromatic/closestpair.jl:39-41,
  src/bichromatic/bichromaticclosestpair.jl:52-139
- neardup: src/neardup.jl:69-256
- hsp_queries: src/hsp.jl:8-33, 77-101
"""
import ctypes as C

import numpy as np

from . import (ExhaustiveSearch, GenericContext, MatrixDatabase, SearchGraph, _as_db, _lib,
               _db_key, _metric, rerank_, searchbatch, searchbatch_knns)

__all__ = ["closestpair", "bichromatic_closestpair", "closestpair_hints", "neardup", "hsp_queries",
           "pairwise"]


def _positions(rows, found):
    """For each r, the position in rows[r] of each id of found[r] (ids distinct within a
    row; entries of found that are 0 map to -1)."""
    nq, k = rows.shape
    off = np.arange(nq, dtype=np.int64)[:, None] << 33
    keys = (off + rows.astype(np.int64)).ravel()
    perm = np.argsort(keys, kind="stable")
    at = np.searchsorted(keys[perm], (off + found.astype(np.int64)).ravel())
    pos = (perm[np.minimum(at, len(perm) - 1)] % k).reshape(nq, -1)
    return np.where(found != 0, pos, -1)


def pairwise(dist, X, ids):
    """(m, k, k) canonical distances between the objects of each row of ids ((m, k),
    1-based, 0 ends a row; pairs with an empty slot are Inf), computed by rerank! on the
    GPU: object ids[i, j] is the query and row i's ids are the candidates (k <= 1024)."""
    X = _as_db(X)
    ids = np.asarray(ids, np.uint32)
    m, k = ids.shape
    if k > 1024:
        raise ValueError("pairwise distances take at most 1024 objects per row")
    out = np.full((m, k, k), np.inf, np.float32)
    qi, qj = np.nonzero(ids > 0)
    if len(qi) == 0:
        return out
    Q = X.rows[ids[qi, qj].astype(np.int64) - 1]                     # one query per object
    cids = np.asfortranarray(ids[qi].T)                               # (k, nq): its row's objects
    cd = np.asfortranarray(np.zeros((k, len(qi)), np.float32))
    rerank_(dist, X, MatrixDatabase.from_rows(Q), cids, cd)
    # rerank! sorts each column by distance: put the distances back by object position
    pos = _positions(ids[qi], cids.T)
    r, t = np.nonzero(pos >= 0)
    out[qi[r], qj[r], pos[r, t]] = cd.T[r, t]
    return out


def bichromatic_closestpair(idxA, ctx, B, min_k=8, samedata=None, seed=0):
    """bichromatic_closestpair(idxA, ctx, B; min_k, samedata) -> (i, j, dist)
    (src/bichromatic/bichromaticclosestpair.jl:52-139): the closest pair (a, b) with a an
    object of idxA and b one of B.  Every object of B is searched with k = min_k; when
    samedata (idxA indexes B itself) a result whose nearest entry is the query itself
    drops it (pop_min!).  The pair of least distance wins, ties to the lowest j.

    Over a SearchGraph with samedata, each object's search is seeded with a random
    neighbor of it and starts with the object itself marked visited, as the reference
    does; the random neighbors come from a numpy generator seeded with `seed`
    (closestpair_hints), not Julia's, so the draws differ from the reference's."""
    B = _as_db(B)
    if samedata is None:
        samedata = idxA.db is B
    if samedata and min_k < 2:
        raise ValueError("min_k must be >= 2 when the index holds B itself")
    n = len(B)
    if n == 0:
        raise ValueError("empty dataset")
    if isinstance(idxA, SearchGraph) and samedata:
        res = searchbatch_knns(idxA, ctx, B, np.full(n, min_k, np.int32),
                               hints=closestpair_hints(idxA, seed),
                               visited=np.arange(1, n + 1, dtype=np.uint32))
        ids = np.stack([r[0] for r in res])
        dists = np.stack([r[1] for r in res])
    else:
        ids, dists = searchbatch(idxA, ctx, B, min_k)
        ids, dists = ids.T, dists.T                                    # (n, k)
    # each object's nearest after pop_min! of itself, then the least distance, lowest j on
    # ties (strict < in object order)
    t = (samedata & (ids[:, 0] == np.arange(1, n + 1, dtype=np.uint32))).astype(np.int64)
    ok = t < ids.shape[1]
    tc = np.minimum(t, ids.shape[1] - 1)
    cid = ids[np.arange(n), tc]
    cd = np.where(ok & (cid != 0), dists[np.arange(n), tc], np.float32(np.inf))
    cd = np.where(np.isnan(cd), np.float32(np.inf), cd)
    if not (cd < np.inf).any():
        return (0, 0, np.float32(np.inf))
    j = int(np.argmin(cd))                                             # first minimum
    return (int(cid[j]), j + 1, cd[j])


def closestpair_hints(G, seed=0):
    """The per-object hint lists bichromatic_search! uses over SearchGraph G with samedata
    (bichromaticclosestpair.jl:111): one random neighbor of each object (numpy generator
    seeded with `seed`), or G's hints for an object without neighbors."""
    rng = np.random.default_rng(seed)
    hints = []
    for i in range(1, len(G) + 1):
        nb = G.adj.neighbors(i)
        hints.append(np.array([nb[rng.integers(len(nb))]] if len(nb) else G.hints, np.uint32))
    return hints


def closestpair(idx, ctx, min_k=8, seed=0):
    """closestpair(idx, ctx; min_k) -> (i, j, dist) (src/bichromatic/closestpair.jl:39-41):
    the closest pair of distinct objects of idx."""
    return bichromatic_closestpair(idx, ctx, idx.db, min_k=min_k, samedata=True, seed=seed)


def _nearest_rows(E, Q, k):
    ids, dists = searchbatch(E, GenericContext(), Q, k)
    return ids[0], dists[0]


def neardup(dist, X, eps, k=8, blocksize=256, filterblocks=True, recall=1.0, parallel_block=128,
            seed=0):
    """neardup(dist, X, ϵ; k, blocksize, filterblocks) with an exact index
    (src/neardup.jl:69-85 with recall = 1.0, then neardup_ :94-171 and neardup_block!
    :186-256), run on the device by ssg_neardup.  X is processed in blocks of blocksize
    objects: a block is searched against the ϵ-net built so far (nearest center within ϵ:
    a near duplicate of it); the rest of the block is filtered in order against the
    block's own new centers (the first center at the least distance, when within ϵ) and
    the new centers join the net.  k only sizes the reference's searches; with an exact
    index their nearest entry does not depend on it.

    With recall < 1 (neardup.jl:71-76) the net is a SearchGraph grown by append_items!
    (index! with OptimizeParameters(MinRecall(recall)) among its callbacks, parallel_block
    and seed as in SearchGraph.build), searched with KnnSorted(k) queues by
    ssg_neardup_graph; costdists then counts the block searches' and filters' evaluations.

    Returns a dict with the reference's fields: idx (an ExhaustiveSearch or SearchGraph
    over the net), map (net position -> object id), nn (each object's covering object,
    1-based; itself for a center), dist (the distance to it), centers (the net, in
    construction order), costdists (distance evaluations of the block searches and
    filters)."""
    X = _as_db(X)
    n = len(X)
    if n and not 1 <= blocksize:
        raise ValueError("blocksize must be >= 1")
    if recall < 1.0:
        return _neardup_graph(dist, X, eps, k, blocksize, filterblocks, recall, parallel_block, seed)
    E = ExhaustiveSearch(dist, X)
    M = np.zeros(max(n, 1), np.uint32)
    L = np.zeros(n, np.int32)
    D = np.zeros(n, np.float32)
    nm = np.zeros(1, np.int64)
    cost = np.zeros(1, np.int64)
    if n:
        _lib.check(_lib.load().ssg_neardup(E._handle(), float(np.float32(eps)), int(blocksize),
                                           int(bool(filterblocks)), M.ctypes.data, nm.ctypes.data,
                                           L.ctypes.data, D.ctypes.data, cost.ctypes.data), "neardup")
    E.close()
    M = M[:int(nm[0])].copy()
    net = ExhaustiveSearch(dist, MatrixDatabase.from_rows(X.rows[M.astype(np.int64) - 1]))
    return {"idx": net, "map": M, "nn": L, "dist": D, "centers": M.copy(),
            "costdists": int(cost[0])}


def _neardup_graph(dist, X, eps, k, blocksize, filterblocks, recall, parallel_block, seed):
    from . import SearchGraph, BeamSearch, _GpuIndex
    n = len(X)
    E = ExhaustiveSearch(dist, X)
    net = SearchGraph.__new__(SearchGraph)
    _GpuIndex.__init__(net, dist, MatrixDatabase.from_rows(np.zeros((0, X.dim), np.float32)))
    net.adj, net.hints, net.algo = None, np.zeros(0, np.uint32), BeamSearch()
    M = np.zeros(max(n, 1), np.uint32)
    L = np.zeros(n, np.int32)
    D = np.zeros(n, np.float32)
    nm = np.zeros(1, np.int64)
    cost = np.zeros(1, np.int64)
    p = _lib.BuildParams(1.3, 2, int(parallel_block), 1.1, 1.5, 256, float(recall), 32, 10, 32,
                         int(seed), 0)
    if n:
        _lib.check(_lib.load().ssg_neardup_graph(E._handle(), net._handle(), float(np.float32(eps)),
                                                 int(blocksize), int(bool(filterblocks)), int(k),
                                                 C.byref(p), M.ctypes.data, nm.ctypes.data,
                                                 L.ctypes.data, D.ctypes.data, cost.ctypes.data),
                   "neardup")
    E.close()
    M = M[:int(nm[0])].copy()
    net.db = MatrixDatabase.from_rows(np.ascontiguousarray(X.rows[M.astype(np.int64) - 1]))
    if len(M):
        net._download()
    return {"idx": net, "map": M, "nn": L, "dist": D, "centers": M.copy(),
            "costdists": int(cost[0])}


_hsp_cache = {}


def hsp_queries(dist, X, Q, knns_ids, knns_dists):
    """hsp_queries(dist, X, Q, knns_ids, knns_dists) -> (hsp_ids, hsp_dists, hsp)
    (src/hsp.jl:8-33, 77-101): each query's candidates ((k, n) columns, 0 ends a column,
    k <= 256) are taken in column order and kept unless one already kept is strictly
    nearer to the candidate than the query is (hsp_should_push).  The kept entries form
    a sorted queue (stable by distance) at the top of each (k, n) output column, then
    0 / Inf; hsp lists each query's (ids, dists).  The filter runs on the GPU
    (ssg_hsp_queries: one wave per query, candidate rows staged in LDS)."""
    X = _as_db(X)
    ids = np.ascontiguousarray(np.asarray(knns_ids, np.uint32).T)     # (n, k) = (k, n) columns
    dd = np.ascontiguousarray(np.asarray(knns_dists, np.float32).T)
    if ids.shape != dd.shape:
        raise ValueError("knns_ids and knns_dists differ in shape")
    n, k = ids.shape
    if n != len(_as_db(Q)):
        raise ValueError("the number of queries is different from the candidate columns")
    key = (_db_key(X), _metric(dist))
    E = _hsp_cache.get(key)
    if E is None or E.db is not X:         # E holds X, so its id cannot be reused
        E = ExhaustiveSearch(dist, X)
        _hsp_cache.clear()
        _hsp_cache[key] = E
    oi = np.zeros((n, k), np.uint32)
    od = np.full((n, k), np.inf, np.float32)
    _lib.check(_lib.load().ssg_hsp_queries(E._handle(), n, k, ids.ctypes.data, dd.ctypes.data,
                                           oi.ctypes.data, od.ctypes.data), "hsp_queries")
    cnt = (oi != 0).sum(1)
    hsp = [(oi[q, :cnt[q]], od[q, :cnt[q]]) for q in range(n)]
    return oi.T, od.T, hsp
