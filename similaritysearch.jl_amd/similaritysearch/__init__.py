"""similaritysearch -- host-side mirror of SimilaritySearch.jl's query API over the
MI355X C ABI (libssg_hip.so, include/ssg.h).

Names follow the reference (`searchbatch`, `searchbatch!` -> `searchbatch_`,
`allknn`, `rerank!` -> `rerank_`, `macrorecall`, `ExhaustiveSearch`,
`ParallelExhaustiveSearch`, `SearchGraph`, `BeamSearch`, `GenericContext`,
`SearchGraphContext`, `MatrixDatabase`, `StaticAdjList`, `Dist.SqL2/L2/NormCosine`).

Layouts are the reference's: a database or query set is a (dim, n) matrix whose
columns are objects (a numpy array of shape (dim, n); its memory is taken as the
(n, dim) row-major run of objects), results are (k, m) matrices with one column
per query (numpy arrays of shape (k, m) in Fortran order), ids are 1-based
uint32 with 0 for an empty slot and +Inf as its distance.  Every search runs on
the GPU; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import SSGError, BeamParams, QueryOpts

__all__ = ["Dist", "SqL2", "L2", "NormCosine", "Cosine", "MatrixDatabase", "StaticAdjList",
           "AdjList", "AdjDict", "BeamSearch", "GenericContext", "SearchGraphContext",
           "ExhaustiveSearch", "ParallelExhaustiveSearch", "SearchGraph", "searchbatch",
           "searchbatch_", "searchbatch_knns", "searchbatch_queues", "allknn", "rerank_",
           "macrorecall", "recallscore", "SSGError", "KnnSorted", "KnnHeap", "knnqueue",
           "unit_rows", "static_adjlist"]


# --------------------------------------------------------------------------- distances
class SqL2:
    """Squared Euclidean distance (src/dist/vecs.jl:86-102)."""
    metric = _lib.SSG_SQL2


class L2:
    """Euclidean distance (src/dist/vecs.jl:52-67)."""
    metric = _lib.SSG_L2


class NormCosine:
    """1 - dot(u, v) for unit vectors (src/dist/cos.jl:39,76-85)."""
    metric = _lib.SSG_NORMCOS


class Cosine:
    """1 - dot(u, v) / (norm(u) norm(v)) (src/dist/cos.jl:103).  Bound as NormCosine over
    copies of the database and the queries whose rows are divided by their norm in FP64
    and rounded to Float32 once; a zero row stays zero (distance 1 to everything, where
    the reference's 0/0 gives NaN).  The reference's own value goes through BLAS dot and
    norm, whose rounding is not reproduced."""
    metric = _lib.SSG_NORMCOS
    normalises = True


class Dist:
    SqL2 = SqL2
    L2 = L2
    NormCosine = NormCosine
    Cosine = Cosine


def unit_rows(R):
    """Rows divided by their Euclidean norm in FP64, rounded to Float32 (zero rows stay
    zero): what Cosine indexes and queries are searched as."""
    R64 = np.asarray(R, np.float64)
    nrm = np.sqrt((R64 * R64).sum(axis=1, keepdims=True))
    return np.ascontiguousarray(np.where(nrm > 0, R64 / np.where(nrm > 0, nrm, 1.0), 0.0), np.float32)


def _rows_for(dist, rows):
    """The rows an index of distance `dist` stores or searches with."""
    return unit_rows(rows) if getattr(dist, "normalises", False) else rows


def _qrows(index, Q):
    """Query rows as the index searches them (normalised for Cosine)."""
    return _rows_for(index.dist, _as_db(Q).rows)


def _metric(dist):
    m = getattr(dist, "metric", None)
    if m is None:
        raise ValueError(f"unsupported distance {dist!r}: use Dist.SqL2, Dist.L2 or Dist.NormCosine")
    return m


# --------------------------------------------------------------------------- data
class MatrixDatabase:
    """Column-major (dim, n) Float32 database (src/db/matrixdatabase.jl:27-59)."""

    def __init__(self, X):
        X = np.asarray(X, dtype=np.float32)
        if X.ndim != 2:
            raise ValueError("MatrixDatabase expects a (dim, n) matrix")
        self.rows = np.ascontiguousarray(X.T)

    @classmethod
    def from_rows(cls, R):
        """Wrap an (n, dim) row-major array of objects without reshaping."""
        db = cls.__new__(cls)
        db.rows = np.ascontiguousarray(R, dtype=np.float32)
        return db

    @property
    def matrix(self):
        return self.rows.T

    @property
    def dim(self):
        return self.rows.shape[1]

    def __len__(self):
        return self.rows.shape[0]

    def __getitem__(self, i):
        """Object i (1-based, like the reference)."""
        return self.rows[i - 1]

    def __setitem__(self, i, v):
        """setindex! of object i (1-based); bumps the version that device copies check."""
        self.rows[i - 1] = np.asarray(v, np.float32)
        self.touch()

    def touch(self):
        """Marks the rows as changed, so cached device copies (rerank_, hsp_queries) are
        uploaded again; call it after writing into `rows` directly."""
        self._version = getattr(self, "_version", 0) + 1


def _db_key(db):
    """What a cached device copy of db depends on: the object, its length, where its rows
    live and its version (bumped by setindex!, push_item!, append_items! and touch())."""
    rows = db.rows
    return (id(db), len(db), rows.ctypes.data if len(db) else 0, getattr(db, "_version", 0))


def _as_db(Q):
    return Q if isinstance(Q, MatrixDatabase) else MatrixDatabase(Q)


class StaticAdjList:
    """Frozen CSR adjacency (src/adj/adjstatic.jl:27-30): node i's neighbours are
    end_point[offset[i-1]+1 : offset[i]] with 1-based ids (offset[0] implicit 0)."""

    def __init__(self, offset, end_point):
        self.offset = np.ascontiguousarray(offset, dtype=np.int64)
        self.end_point = np.ascontiguousarray(end_point, dtype=np.uint32)

    @classmethod
    def from_csr(cls, offs, adj):
        """From an (n+1)-entry exclusive prefix and 1-based ids."""
        offs = np.asarray(offs, dtype=np.int64)
        return cls(offs[1:], adj)

    def __len__(self):
        return len(self.offset)

    def neighbors(self, i):
        s = 0 if i == 1 else int(self.offset[i - 2])
        return self.end_point[s:int(self.offset[i - 1])]

    def neighbors_length(self, i):
        return int(self.offset[i - 1]) - (0 if i == 1 else int(self.offset[i - 2]))


class AdjList:
    """AdjList (src/adj/adjlist.jl): one neighbour vector per node, 1-based ids; a node
    never assigned a list has none.  SearchGraph freezes it into a StaticAdjList."""

    def __init__(self, lists):
        self.end_point = [None if x is None else np.asarray(x, np.uint32) for x in lists]

    def __len__(self):
        return len(self.end_point)

    def neighbors(self, i):
        return self.end_point[i - 1]

    def neighbors_length(self, i):
        x = self.end_point[i - 1]
        return 0 if x is None else len(x)


class AdjDict:
    """AdjDict (src/adj/adjdict.jl): node id -> neighbour vector; absent nodes have none.
    n is the number of nodes (ids 1..n)."""

    def __init__(self, d, n):
        self.end_point = {int(k): np.asarray(v, np.uint32) for k, v in d.items()}
        self.n = int(n)

    def __len__(self):
        return self.n

    def neighbors(self, i):
        return self.end_point.get(int(i))

    def neighbors_length(self, i):
        x = self.end_point.get(int(i))
        return 0 if x is None else len(x)


def static_adjlist(adj):
    """A StaticAdjList holding the same lists as adj (StaticAdjList, AdjList or AdjDict),
    in node order, each list in its own order (StaticAdjList(adj), src/adj/adjstatic.jl:
    69-120)."""
    if isinstance(adj, StaticAdjList):
        return adj
    if not isinstance(adj, (AdjList, AdjDict)):
        raise TypeError("adj must be a StaticAdjList, AdjList or AdjDict")
    n = len(adj)
    lens = np.array([adj.neighbors_length(i) for i in range(1, n + 1)], np.int64)
    ep = np.zeros(int(lens.sum()), np.uint32)
    pos = 0
    for i in range(1, n + 1):
        if lens[i - 1]:
            ep[pos:pos + lens[i - 1]] = adj.neighbors(i)
            pos += int(lens[i - 1])
    return StaticAdjList(np.cumsum(lens), ep)


# --------------------------------------------------------------------------- search params
class BeamSearch:
    """BeamSearch(; bsize=4, Δ=1.0, maxvisits=10^6) (src/searchgraph/SearchGraph.jl:94-100)."""

    def __init__(self, bsize=4, delta=1.0, maxvisits=10**6, **kw):
        if "Δ" in kw:
            delta = kw.pop("Δ")
        if kw:
            raise TypeError(f"unknown arguments {sorted(kw)}")
        self.bsize = int(bsize)
        self.delta = float(np.float32(delta))
        self.maxvisits = int(maxvisits)

    def __repr__(self):
        return f"BeamSearch(bsize={self.bsize}, Δ={self.delta}, maxvisits={self.maxvisits})"


class GenericContext:
    """Context of exhaustive indexes.  costdists holds the per-query distance
    evaluations of the last batch call (src/SimilaritySearch.jl:187-258)."""

    def __init__(self, verbose=False, precision="f32", ncand=0):
        self.verbose = verbose
        self.costdists = np.zeros(0, np.int64)
        self.costblocks = np.zeros(0, np.int64)
        if precision not in ("f32", "bf16", "fp8"):
            raise ValueError("precision must be 'f32', 'bf16' or 'fp8'")
        # precision form of the exhaustive indexes (no reference counterpart): a BF16 or
        # FP8 candidate pass keeping max(ncand, 4k) candidates per query, then rerank!
        self.precision = precision
        self.ncand = int(ncand)

    def exhaustive_ncand(self, k):
        return min(256, max(k, self.ncand if self.ncand > 0 else 4 * k))


_PREC_CODE = {"bf16": 1, "fp8": 2}  # SSG_BF16, SSG_FP8


class SearchGraphContext(GenericContext):
    """Context of SearchGraph; beamcap is the number of rows of its beam buffers,
    the cap getbeam applies to bsize (src/searchgraph/context.jl:119-126,220-225)."""

    def __init__(self, beamcap=32, verbose=False, precision="f32", rerank=0, expand=1):
        super().__init__(verbose)
        if not 1 <= beamcap <= 256:
            raise ValueError("beamcap must be in 1..256")
        if precision not in ("f32", "bf16"):
            raise ValueError("precision must be 'f32' or 'bf16'")
        if not 1 <= expand <= 4 or (expand > 1 and precision != "bf16"):
            raise ValueError("expand must be 1..4, and above 1 only with precision 'bf16'")
        self.beamcap = int(beamcap)
        # precision form (no reference counterpart): BF16 gathers during the search and an
        # exact rerank of the `rerank` best candidates
        self.precision = precision
        self.rerank = int(rerank)
        # wide steps of the BF16 form: beam entries popped and expanded together
        self.expand = int(expand)


# --------------------------------------------------------------------------- indexes
class _GpuIndex:
    def __init__(self, dist, db, device=0):
        self.dist = dist
        self.db = _as_db(db)
        self.device = device
        self._h = None

    def _handle(self):
        if self._h is None:
            if self.db.rows.dtype != np.float32 or not self.db.rows.flags.c_contiguous:
                raise TypeError("the GPU indexes take a contiguous Float32 database")
            L = _lib.load()
            h = C.c_void_p()
            _lib.check(L.ssg_create(C.byref(h), self.device), "ssg_create")
            self._h = h
            X = _rows_for(self.dist, self.db.rows)
            _lib.check(L.ssg_upload_db(h, X.ctypes.data, X.shape[1], X.shape[0],
                                       _metric(self.dist)), "ssg_upload_db")
            self._after_upload()
        return self._h

    def _after_upload(self):
        pass

    def close(self):
        if self._h is not None:
            _lib.load().ssg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.db)

    def database(self, i=None):
        return self.db if i is None else self.db[i]

    def distance(self):
        return self.dist


class ExhaustiveSearch(_GpuIndex):
    """Exact k-NN index (src/exact/sequential-exhaustive.jl)."""

    def __init__(self, dist=None, db=None, device=0):
        super().__init__(dist if dist is not None else SqL2, db, device)

    def append_items(self, ctx, U):
        """append_items!(index, ctx, U) (src/exact/sequential-exhaustive.jl:40-43): U's
        objects get ids n+1..n+len(U).  The host rows live in a buffer whose capacity
        doubles, and an uploaded index appends only the new rows (ssg_append_db)."""
        U = _as_db(U)
        m = len(U)
        if m == 0:
            return self
        if U.dim != self.db.dim:
            raise ValueError("appended objects differ in dimension")
        rows = np.ascontiguousarray(U.rows, np.float32)
        n = len(self.db)
        buf = getattr(self, "_buf", None)
        if buf is None or buf.shape[0] < n + m or not np.shares_memory(buf, self.db.rows):
            cap = max(n + m, 2 * (0 if buf is None else buf.shape[0]), 64)
            nb = np.empty((cap, self.db.dim), np.float32)
            nb[:n] = self.db.rows
            self._buf = buf = nb
        buf[n:n + m] = rows
        self.db = MatrixDatabase.from_rows(buf[:n + m])
        if self._h is not None:
            up = _rows_for(self.dist, rows)
            _lib.check(_lib.load().ssg_append_db(self._h, up.ctypes.data, m), "append_items!")
        return self

    def push_item(self, ctx, u):
        """push_item!(index, ctx, u) (src/exact/sequential-exhaustive.jl:33-37)."""
        return self.append_items(ctx, MatrixDatabase.from_rows(np.asarray(u, np.float32)[None, :]))


class ParallelExhaustiveSearch(ExhaustiveSearch):
    """Same results as ExhaustiveSearch (src/exact/parallel-exhaustive.jl:64-102);
    on the GPU the DB split is the MFMA kernel's slice axis."""


class SearchGraph(_GpuIndex):
    """A frozen SearchGraph (src/searchgraph/SearchGraph.jl:147-158)."""

    def __init__(self, dist, db, adj, hints=(), algo=None, device=0):
        super().__init__(dist, db, device)
        adj = static_adjlist(adj)
        if len(adj) != len(self.db):
            raise ValueError("graph and database sizes differ")
        self.adj = adj
        self.hints = np.ascontiguousarray(hints, dtype=np.uint32)
        self.algo = algo if algo is not None else BeamSearch()

    def _after_upload(self):
        a = self.adj
        if a is None:  # being built by index!
            return
        _lib.check(_lib.load().ssg_upload_graph(self._h, a.offset.ctypes.data, len(a),
                                               a.end_point.ctypes.data, len(a.end_point),
                                               self.hints.ctypes.data, len(self.hints)),
                   "ssg_upload_graph")

    def beam_params(self, ctx, algo=None):
        b = algo or self.algo
        prec = 1 if getattr(ctx, "precision", "f32") == "bf16" else 0
        return BeamParams(b.bsize, b.delta, b.maxvisits, getattr(ctx, "beamcap", 32), prec,
                          getattr(ctx, "rerank", 0), getattr(ctx, "expand", 1))

    @classmethod
    def build(cls, dist, db, ctx=None, parallel_block=128, neighborhood_logbase=1.3,
              neighborhood_minsize=2, hints_logbase=1.1, logbase_callback=1.5,
              starting_callback=256, minrecall=0.97, numqueries=32, ksearch=10, seed=0,
              device=0, verbose=0):
        """SearchGraph(dist, db) followed by index!(G, ctx) (src/searchgraph/insertions.jl:94-113)
        with the reference's default context knobs; the graph is built on the GPU."""
        G = cls.__new__(cls)
        _GpuIndex.__init__(G, dist, db, device)
        G.adj, G.hints, G.algo = None, np.zeros(0, np.uint32), BeamSearch()
        ctx = ctx or SearchGraphContext()
        h = G._handle()
        p = _lib.BuildParams(neighborhood_logbase, neighborhood_minsize, parallel_block,
                             hints_logbase, logbase_callback, starting_callback, minrecall,
                             numqueries, ksearch, ctx.beamcap, seed, verbose)
        L = _lib.load()
        _lib.check(L.ssg_build_graph(h, C.byref(p)), "index!")
        G._download()
        return G

    def _download(self):
        """The device graph, its hints and BeamSearch copied into adj / hints / algo."""
        L = _lib.load()
        h = self._handle()
        nnz, nh, bp = C.c_int64(), C.c_int64(), BeamParams()
        _lib.check(L.ssg_graph_info(h, C.byref(nnz), C.byref(nh), C.byref(bp)), "graph info")
        n = len(self.db)
        offset = np.zeros(n, np.int64)
        end_point = np.zeros(max(nnz.value, 1), np.uint32)
        hints = np.zeros(max(nh.value, 1), np.uint32)
        _lib.check(L.ssg_download_graph(h, offset.ctypes.data, end_point.ctypes.data,
                                        hints.ctypes.data), "graph download")
        self.adj = StaticAdjList(offset, end_point[:nnz.value])
        self.hints = hints[:nh.value].copy()
        self.algo = BeamSearch(bp.bsize, bp.delta, bp.maxvisits)

    def append_items(self, ctx, U):
        """append_items!(G, ctx, U) (src/searchgraph/insertions.jl:27-34) for a graph that
        SearchGraph.build made: U's objects get ids n+1..n+len(U) and index! continues over
        them on the device with the build's parameters (ssg_append_items_graph)."""
        U = _as_db(U)
        m = len(U)
        if m == 0:
            return self
        if U.dim != self.db.dim:
            raise ValueError("appended objects differ in dimension")
        rows = np.ascontiguousarray(U.rows, np.float32)
        up = _rows_for(self.dist, rows)
        _lib.check(_lib.load().ssg_append_items_graph(self._handle(), up.ctypes.data, m),
                   "append_items!")
        self.db = MatrixDatabase.from_rows(np.concatenate([self.db.rows, rows]))
        self._download()
        return self

    def set_launch(self, slots_per_cu=32, rounds=8, vbuckets=-1):
        """Launch shape of this index's searches (ssg_set_launch; results do not depend on it)."""
        _lib.check(_lib.load().ssg_set_launch(self._handle(), C.byref(_lib.LaunchParams(
            int(slots_per_cu), int(rounds), int(vbuckets)))), "set launch")

    def launch(self):
        lp = _lib.LaunchParams()
        _lib.check(_lib.load().ssg_get_launch(self._handle(), C.byref(lp)), "get launch")
        return lp.slots_per_cu, lp.rounds, lp.vbuckets

    def optimize_launch_dev(self, bp, dQ, m, k, kstride, gold, minrecall, qo=None):
        """ssg_optimize_launch over m device-resident queries (qo: device QueryOpts) with
        host gold ids (m, kg): installs the fastest launch shape among the recall-changing
        knobs that reach minrecall and writes the chosen rerank/expand into bp.  Returns
        ((slots_per_cu, rounds, vbuckets), recall, ms)."""
        gold = np.ascontiguousarray(gold, np.uint32)
        lp = _lib.LaunchParams()
        rec, ms = C.c_double(), C.c_float()
        _lib.check(_lib.load().ssg_optimize_launch(
            self._handle(), dQ, m, k, kstride, C.byref(bp), None if qo is None else C.byref(qo),
            gold.ctypes.data, gold.shape[1], float(minrecall), C.byref(lp), C.byref(rec),
            C.byref(ms)), "optimize launch")
        return (lp.slots_per_cu, lp.rounds, lp.vbuckets), rec.value, ms.value

    def optimize(self, ctx, minrecall, queries, ksearch=10):
        """optimize_index!(G, ctx, MinRecall(r); queries, ksearch) (src/opt.jl:220-317);
        installs and returns the chosen BeamSearch and the sample recall."""
        Q = _as_db(queries)
        bp = self.beam_params(ctx)
        rec = C.c_double()
        rows = _qrows(self, Q)
        _lib.check(_lib.load().ssg_optimize_minrecall(self._handle(), rows.ctypes.data, len(Q),
                                                      ksearch, minrecall, C.byref(bp),
                                                      C.byref(rec)), "optimize_index!")
        self.algo = BeamSearch(bp.bsize, bp.delta, bp.maxvisits)
        return self.algo, rec.value


# --------------------------------------------------------------------------- batch API
def _out(m, k):
    ids = np.zeros((m, k), np.uint32)
    dists = np.full((m, k), np.inf, np.float32)
    return ids, dists


def searchbatch(index, ctx, Q, k, sorted=True):
    """searchbatch(index, ctx, Q, k) -> (ids, dists), both (k, m)
    (src/SimilaritySearch.jl:309-313)."""
    Q = _as_db(Q)
    m = len(Q)
    if m == 0:
        raise ValueError("empty set of queries")
    ids, dists = _out(m, k)
    searchbatch_(index, ctx, Q, ids.T, dists.T, sorted=sorted)
    return ids.T, dists.T


def searchbatch_(index, ctx, Q, ids, dists, sorted=False):
    """searchbatch!(index, ctx, Q, ids, dists) in place on (k, m) Fortran-ordered
    outputs (src/SimilaritySearch.jl:331-348)."""
    Q = _as_db(Q)
    m = len(Q)
    if m == 0:
        raise ValueError("empty set of queries")
    if ids.shape[1] != m or dists.shape != ids.shape:
        raise ValueError("the number of queries is different from the given output containers")
    if not (ids.T.flags.c_contiguous and dists.T.flags.c_contiguous):
        raise ValueError("outputs must be Fortran-ordered (k, m) arrays")
    if ids.dtype != np.uint32 or dists.dtype != np.float32:
        raise TypeError("ids must be UInt32 and dists Float32")
    k = ids.shape[0]
    L = _lib.load()
    h = index._handle()
    rows = _qrows(index, Q)
    if rows.shape[1] != index.db.dim:
        raise ValueError("query dimension differs from the database")
    if isinstance(index, SearchGraph):
        cd = np.zeros(m, np.int64)
        cb = np.zeros(m, np.int64)
        bp = index.beam_params(ctx)
        _lib.check(L.ssg_beamsearch(h, rows.ctypes.data, m, k, k, C.byref(bp), None,
                                    ids.T.ctypes.data, dists.T.ctypes.data, cd.ctypes.data,
                                    cb.ctypes.data), "searchbatch!")
        ctx.costdists, ctx.costblocks = cd, cb
    elif getattr(ctx, "precision", "f32") in _PREC_CODE:
        _lib.check(L.ssg_exhaustive_knn_prec(h, rows.ctypes.data, m, k, _PREC_CODE[ctx.precision],
                                             ctx.exhaustive_ncand(k), ids.T.ctypes.data,
                                             dists.T.ctypes.data, None, None), "searchbatch!")
        ctx.costdists = np.full(m, len(index), np.int64)
        ctx.costblocks = np.zeros(m, np.int64)
    else:
        _lib.check(L.ssg_exhaustive_knn(h, rows.ctypes.data, m, k, ids.T.ctypes.data,
                                        dists.T.ctypes.data), "searchbatch!")
        ctx.costdists = np.full(m, len(index), np.int64)
        ctx.costblocks = np.zeros(m, np.int64)
    return ids, dists


def searchbatch_knns(index, ctx, Q, ks, bsizes=None, hints=None, visited=None):
    """Queue-vector form searchbatch!(index, ctx, Q, knns) (src/SimilaritySearch.jl:384-398)
    with per-query k (ks), optional per-query beam sizes and per-query hint lists
    (1-based; `hints` is a list of sequences).  `visited` gives per query a 1-based object
    marked visited before its search starts (0 = none), as bichromatic_search! does with
    the query's own vertex.  Returns a list of (ids, dists) pairs."""
    if not isinstance(index, SearchGraph):
        raise TypeError("per-query queues are supported over SearchGraph")
    Q = _as_db(Q)
    m = len(Q)
    if m == 0:
        raise ValueError("empty set of queries")
    ks = np.ascontiguousarray(ks, np.int32)
    if len(ks) != m:
        raise ValueError("the number of queries is different from the given output containers")
    kstride = int(ks.max())
    ids, dists = _out(m, kstride)
    cd = np.zeros(m, np.int64)
    cb = np.zeros(m, np.int64)
    keep = [ks]
    qo = QueryOpts(ks.ctypes.data, None, None, None, 0)
    if visited is not None:
        v = np.ascontiguousarray(visited, np.uint32)
        if len(v) != m:
            raise ValueError("visited needs one entry per query")
        keep.append(v)
        qo.visited = v.ctypes.data
    if bsizes is not None:
        b = np.ascontiguousarray(bsizes, np.int32)
        keep.append(b)
        qo.bsize = b.ctypes.data
    if hints is not None:
        stride = max(1, max(len(x) for x in hints))
        H = np.zeros((m, stride), np.uint32)
        cnt = np.zeros(m, np.int32)
        for j, x in enumerate(hints):
            H[j, :len(x)] = x
            cnt[j] = len(x)
        keep += [H, cnt]
        qo.hints, qo.nhints, qo.hint_stride = H.ctypes.data, cnt.ctypes.data, stride
    bp = index.beam_params(ctx)
    rows = _qrows(index, Q)
    _lib.check(_lib.load().ssg_beamsearch(index._handle(), rows.ctypes.data, m, kstride,
                                          kstride, C.byref(bp), C.byref(qo), ids.ctypes.data,
                                          dists.ctypes.data, cd.ctypes.data, cb.ctypes.data),
               "searchbatch!")
    ctx.costdists, ctx.costblocks = cd, cb
    return [(ids[j, :ks[j]].copy(), dists[j, :ks[j]].copy()) for j in range(m)]


def searchbatch_queues(index, ctx, Q, knns):
    """searchbatch!(index, ctx, Q, knns) with a vector of KnnSorted / KnnHeap queues
    (src/SimilaritySearch.jl:384-398): query j's first maxlength(knns[j]) results in
    (distance, id) order are pushed into knns[j] with the queue's own push_item!
    (similaritysearch.pqueue).  Over a SearchGraph every query runs with its own k in one
    launch; over an exhaustive index the batch runs once at the largest k."""
    from .pqueue import fill
    Q = _as_db(Q)
    m = len(Q)
    if m == 0:
        raise ValueError("empty set of queries")
    if len(knns) != m:
        raise ValueError("the number of queries is different from the given output containers")
    ks = np.array([r.maxlength() for r in knns], np.int32)
    if (ks < 1).any():
        raise ValueError("every queue needs room for one item")
    if isinstance(index, SearchGraph):
        runs = searchbatch_knns(index, ctx, Q, ks)
    else:
        kmax = int(ks.max())
        ids, dists = _out(m, kmax)
        searchbatch_(index, ctx, Q, ids.T, dists.T)
        runs = [(ids[j, :ks[j]], dists[j, :ks[j]]) for j in range(m)]
    for r, (i, d) in zip(knns, runs):
        fill(r, i, d)
    return knns


def allknn(index, ctx, k, sort=True):
    """allknn(index, ctx, k) -> (ids, dists), both (k, n) (src/allknn.jl:39-91)."""
    n = len(index)
    if not 0 < k <= n:
        raise ValueError("allknn needs 0 < k <= n")
    ids, dists = _out(n, k)
    L = _lib.load()
    h = index._handle()
    if isinstance(index, SearchGraph):
        bp = index.beam_params(ctx)
        _lib.check(L.ssg_allknn_graph(h, k, C.byref(bp), ids.ctypes.data, dists.ctypes.data),
                   "allknn")
    elif getattr(ctx, "precision", "f32") in _PREC_CODE:
        _lib.check(L.ssg_allknn_exhaustive_prec(h, k, _PREC_CODE[ctx.precision], ctx.exhaustive_ncand(k),
                                                ids.ctypes.data, dists.ctypes.data, None, None), "allknn")
    else:
        _lib.check(L.ssg_allknn_exhaustive(h, k, ids.ctypes.data, dists.ctypes.data), "allknn")
    return ids.T, dists.T


def exhaustive_candidates(index, k, ncand, Q=None, precision="bf16"):
    """Precision form of an exhaustive index with its candidate queue exposed (no reference
    counterpart): returns (ids, dists, cand_ids, cand_dists), all (rows, m) with m the
    number of queries (every database object when Q is None, as allknn).  cand_* hold
    the ncand candidates of the BF16 or FP8 pass in exact (distance, id) order and their
    approximate distances; ids/dists are the exact top k among them."""
    if not isinstance(index, ExhaustiveSearch):
        raise TypeError("the precision form serves exhaustive indexes")
    if precision not in _PREC_CODE:
        raise ValueError("precision must be 'bf16' or 'fp8'")
    pc = _PREC_CODE[precision]
    L = _lib.load()
    h = index._handle()
    if Q is None:
        m = len(index)
        if not 0 < k <= m:
            raise ValueError("allknn needs 0 < k <= n")
    else:
        Q = _as_db(Q)
        m = len(Q)
        if m == 0:
            raise ValueError("empty set of queries")
    ids, dists = _out(m, k)
    ci, cd = _out(m, ncand)
    if Q is None:
        _lib.check(L.ssg_allknn_exhaustive_prec(h, k, pc, ncand, ids.ctypes.data, dists.ctypes.data,
                                                ci.ctypes.data, cd.ctypes.data), "allknn")
    else:
        if Q.dim != index.db.dim:
            raise ValueError("query dimension differs from the database")
        rows = _qrows(index, Q)
        _lib.check(L.ssg_exhaustive_knn_prec(h, rows.ctypes.data, m, k, pc, ncand, ids.ctypes.data,
                                             dists.ctypes.data, ci.ctypes.data, cd.ctypes.data),
                   "searchbatch!")
    return ids.T, dists.T, ci.T, cd.T


_rerank_cache = {}


def rerank_(dist, db, Q, ids, dists):
    """rerank!(dist, db, queries, ids, dists) batch form, in place (src/rerank.jl:77-86)."""
    Q = _as_db(Q)
    key = (_db_key(db), _metric(dist))
    hit = _rerank_cache.get(key)
    if hit is not None and hit[0] is db:   # the entry holds db, so its id cannot be reused
        idx = hit[1]
    else:
        idx = ExhaustiveSearch(dist, db)
        _rerank_cache.clear()              # one uploaded database at a time
        _rerank_cache[key] = (db, idx)
    m = len(Q)
    k = ids.shape[0]
    if not (ids.T.flags.c_contiguous and dists.T.flags.c_contiguous):
        raise ValueError("outputs must be Fortran-ordered (k, m) arrays")
    rows = _qrows(idx, Q)
    _lib.check(_lib.load().ssg_rerank(idx._handle(), rows.ctypes.data, m, k, ids.T.ctypes.data,
                                      dists.T.ctypes.data), "rerank!")
    return ids, dists


def fft(dist, X, k, start=1, device=0):
    """fft(dist, X, k; start) farthest-first k-centers on the GPU (src/kcenters/fft.jl:53-87).
    Returns (centers, nn, dists, eps) with 1-based ids; eps[i] is the ε recorded when
    center i was added (Inf for the first)."""
    E = X if isinstance(X, _GpuIndex) else ExhaustiveSearch(dist, X, device)
    n = len(E)
    centers = np.zeros(k, np.uint32)
    nn = np.zeros(n, np.uint32)
    nnd = np.zeros(n, np.float32)
    eps = np.zeros(k, np.float32)
    _lib.check(_lib.load().ssg_fft(E._handle(), k, start, centers.ctypes.data, nn.ctypes.data,
                                   nnd.ctypes.data, eps.ctypes.data), "fft")
    return centers, nn, nnd, eps


# --------------------------------------------------------------------------- metrics
def recallscore(gold, res):
    """|Set(gold) ∩ Set(res)| / length(gold) (src/perf.jl:31-33)."""
    gold = np.asarray(gold)
    return len(set(gold.tolist()) & set(np.asarray(res).tolist())) / len(gold)


def macrorecall(gold, res, k=None):
    """Mean per-column recallscore of (k, m) id matrices (src/perf.jl:71-79)."""
    gold = np.asarray(gold)
    res = np.asarray(res)
    k = gold.shape[0] if k is None else k
    return float(np.mean([recallscore(gold[:k, j], res[:k, j]) for j in range(gold.shape[1])]))


from .consumers import (bichromatic_closestpair, closestpair, hsp_queries, neardup,  # noqa: E402
                        pairwise)

__all__ += ["closestpair", "bichromatic_closestpair", "neardup", "hsp_queries", "pairwise"]

from .pqueue import KnnHeap, KnnSorted, knnqueue  # noqa: E402
from .storage import (MMapMatrixDatabase, load_index, load_results, read_matrix,  # noqa: E402
                      save_index, split_pathkey, write_results)

__all__ += ["MMapMatrixDatabase", "save_index", "load_index", "read_matrix", "split_pathkey",
            "write_results", "load_results"]
