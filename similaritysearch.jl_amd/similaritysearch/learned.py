"""Learned entry points for SearchGraph beam search (no reference counterpart).

Pipeline, all on the GPU:
  1. landmarks: `fft` k-centers of the database (src/kcenters/fft.jl:53-87), start=1;
     every object's nearest landmark comes with it.
  2. traces: for each training query, the exact top-10 (exhaustive search) gives the
     label landmark (the landmark nearest to the query's true nearest neighbour) and,
     by running the search itself from that landmark with growing beam sizes, the
     smallest beam-size bucket whose result equals the exact top-10.
  3. an MLP (ssg_mlp, hand-written HIP forward/backward/Adam kernels) learns
     query -> (landmark, bucket).
  4. search: inference gives the top `nhints` landmarks as per-query hints and the
     bucket's beam size; queries whose top-landmark probability is below `min_conf`
     fall back to the plain search (SearchGraph.hints and the plain beam size).
     The search itself runs in the BF16 precision form with an exact rerank.
  5. tune: the knobs (quantile, rerank length, beam-size cap, delta) and the fallback are
     chosen on tuning queries against exact ground truth.  Each query's recall is
     predicted from its landmark confidence by an isotonic (monotone) fit of the recall
     the learned search reached on the tuning queries; queries whose predicted recall is
     below the target take the plain search with the smallest beam size whose own recall
     reaches the target.  The cheapest setting (mean distance evaluations) whose overall
     tuning recall clears the target with a margin wins: the recall-targeted cost of
     optimize_index! (src/opt.jl:220-317, MinRecall) over these knobs.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import BeamParams, QueryOpts
from .device import DeviceArray, Stream

DEFAULT_BUCKETS = (2, 4, 8, 16, 32, 64, 128, 256)


class Mlp:
    """D -> H1 -> H2 -> (L + NB) MLP on the GPU (include/ssg.h, ssg_mlp_*)."""

    def __init__(self, D, H1, H2, L, NB, seed=0, device=0):
        self.shape = (D, H1, H2, L, NB)
        self._p = C.c_void_p()
        _lib.check(_lib.load().ssg_mlp_create(C.byref(self._p), device, D, H1, H2, L, NB, seed),
                   "ssg_mlp_create")

    def _dims(self, layer):
        D, H1, H2, L, NB = self.shape
        return [(H1, D), (H2, H1), (L + NB, H2)][layer]

    def get(self, layer):
        o, i = self._dims(layer)
        W = np.zeros((o, i), np.float32)
        b = np.zeros(o, np.float32)
        _lib.check(_lib.load().ssg_mlp_get(self._p, layer, W.ctypes.data, b.ctypes.data), "get")
        return W, b

    def set(self, layer, W, b):
        W = np.ascontiguousarray(W, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        _lib.check(_lib.load().ssg_mlp_set(self._p, layer, W.ctypes.data, b.ctypes.data), "set")

    def grads(self, layer):
        o, i = self._dims(layer)
        gW = np.zeros((o, i), np.float32)
        gb = np.zeros(o, np.float32)
        _lib.check(_lib.load().ssg_mlp_grad(self._p, layer, gW.ctypes.data, gb.ctypes.data), "grad")
        return gW, gb

    def step(self, X, yl, yb, bucket_weight=1.0, lr=1e-3, update=True):
        X = np.ascontiguousarray(X, np.float32)
        yl = np.ascontiguousarray(yl, np.int32)
        yb = np.ascontiguousarray(yb, np.int32)
        loss = C.c_float()
        _lib.check(_lib.load().ssg_mlp_step(self._p, X.ctypes.data, yl.ctypes.data, yb.ctypes.data,
                                            len(X), bucket_weight, lr, int(update), C.byref(loss)),
                   "mlp step")
        return loss.value

    def train_dev(self, dX, dyl, dyb, N, epochs, batch, bucket_weight=1.0, lr=1e-3, seed=0):
        losses = np.zeros(epochs, np.float32)
        _lib.check(_lib.load().ssg_mlp_train_dev(self._p, dX, dyl, dyb, N, epochs, batch,
                                                 bucket_weight, lr, seed, losses.ctypes.data),
                   "mlp train")
        return losses

    def close(self):
        if self._p:
            _lib.load().ssg_mlp_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedGraph:
    """A recorded learned-search sequence; launch() replays it on a stream."""

    def __init__(self, g):
        self._g = g

    def launch(self, stream):
        _lib.check(_lib.load().ssg_graph_launch(self._g, stream.handle), "graph launch")

    def close(self):
        if self._g:
            _lib.load().ssg_graph_destroy(self._g)
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedSearch:
    """Learned entry points + BF16 beam search over a SearchGraph `G`."""

    def __init__(self, G, nlandmarks=4096, hidden=256, buckets=DEFAULT_BUCKETS, seed=9):
        self.G = G
        self.L = int(nlandmarks)
        self.hidden = int(hidden)
        self.buckets = tuple(int(b) for b in buckets)
        self.seed = seed
        self.mlp = None
        self.landmarks = None

    # ---------------------------------------------------------------- traces
    def landmark_labels(self):
        from . import fft
        centers, nn, _, _ = fft(self.G.dist, self.G, self.L, start=1)
        self.landmarks = centers
        pos = np.full(len(self.G) + 1, -1, np.int64)
        pos[centers.astype(np.int64)] = np.arange(self.L)
        self.object_landmark = pos[nn.astype(np.int64)]  # landmark index of each object
        return centers

    def traces(self, Qtrain, ctx, k=10, rerank=20, batch=50_000, min_hits=None):
        """Labels (landmark index, bucket index) for each training query row: the bucket is
        the smallest beam size whose search from the label landmark finds at least
        `min_hits` (default k) of the exact k nearest neighbours."""
        from . import _rows_for
        min_hits = k if min_hits is None else int(min_hits)
        G = self.G
        Qtrain = _rows_for(G.dist, np.ascontiguousarray(Qtrain, np.float32))
        h = G._handle()
        L = _lib.load()
        m = len(Qtrain)
        yl = np.zeros(m, np.int32)
        yb = np.full(m, len(self.buckets) - 1, np.int32)
        found = np.zeros(m, bool)
        for s in range(0, m, batch):
            e = min(m, s + batch)
            Qb = np.ascontiguousarray(Qtrain[s:e], np.float32)
            gi = np.zeros((e - s, k), np.uint32)
            gd = np.zeros((e - s, k), np.float32)
            _lib.check(L.ssg_exhaustive_knn(h, Qb.ctypes.data, e - s, k, gi.ctypes.data,
                                            gd.ctypes.data), "exhaustive")
            yl[s:e] = self.object_landmark[gi[:, 0].astype(np.int64) - 1]
            hints = self.landmarks[yl[s:e]].astype(np.uint32).reshape(-1, 1).copy()
            cnt = np.ones(e - s, np.int32)
            gold = np.sort(gi, axis=1)
            fnd = np.zeros(e - s, bool)
            for bi, b in enumerate(self.buckets):
                bq = np.full(e - s, b, np.int32)
                qo = QueryOpts(None, bq.ctypes.data, hints.ctypes.data, cnt.ctypes.data, 1)
                bp = BeamParams(b, G.algo.delta, 10**6, max(ctx.beamcap, b), 1, rerank)
                ids = np.zeros((e - s, k), np.uint32)
                ds = np.zeros((e - s, k), np.float32)
                _lib.check(L.ssg_beamsearch(h, Qb.ctypes.data, e - s, k, k, C.byref(bp), C.byref(qo),
                                            ids.ctypes.data, ds.ctypes.data, None, None), "trace search")
                hits = (ids[:, :, None] == gold[:, None, :]).any(axis=2).sum(axis=1)
                ok = (hits >= min_hits) & ~fnd
                yb[s:e][ok] = bi
                fnd |= ok
            found[s:e] = fnd
        return yl, yb, found

    # ---------------------------------------------------------------- training
    def fit(self, Qtrain, ctx, epochs=6, batch=1024, lr=2e-3, bucket_weight=0.5, verbose=None):
        if self.landmarks is None:
            self.landmark_labels()
        yl, yb, found = self.traces(Qtrain, ctx)
        if verbose:
            verbose(f"traces: {len(Qtrain)} queries, exact top-10 reached for {found.mean():.3f}, "
                    f"bucket histogram {np.bincount(yb, minlength=len(self.buckets)).tolist()}")
        from . import _rows_for
        D = Qtrain.shape[1]
        self.mlp = Mlp(D, self.hidden, self.hidden, self.L, len(self.buckets), seed=self.seed)
        dX = DeviceArray.from_numpy(_rows_for(self.G.dist, np.ascontiguousarray(Qtrain, np.float32)))
        dyl = DeviceArray.from_numpy(yl)
        dyb = DeviceArray.from_numpy(yb)
        losses = self.mlp.train_dev(dX.ptr, dyl.ptr, dyb.ptr, len(Qtrain), epochs, batch,
                                    bucket_weight, lr, self.seed)
        self.dlandmarks = DeviceArray.from_numpy(self.landmarks)
        self.train_labels = (yl, yb)
        return losses

    # ---------------------------------------------------------------- search
    def table(self, shift=0):
        """Beam size of each bucket: buckets[b + shift], clamped to the bucket range."""
        nb = len(self.buckets)
        return np.array([self.buckets[min(nb - 1, max(0, b + shift))] for b in range(nb)], np.int32)

    def search_dev(self, ctx, dQ, m, k, dI, dD, nhints=1, min_conf=0.0, shift=0, rerank=20,
                   fallback_bsize=None, stream=None, dcd=None, dcb=None, expand=1, quantile=0.0,
                   delta=None):
        """Device-pointer learned search.  Each query's bucket is the most probable one
        (quantile 0) or the smallest whose cumulative probability reaches `quantile`;
        bucket b maps to beam size buckets[b + shift]; `expand` beam entries are popped per
        step (1: the reference step); `delta` defaults to the graph's tuned one."""
        G = self.G
        table = self.table(shift)
        fb = fallback_bsize or G.algo.bsize
        d = G.algo.delta if delta is None else float(delta)
        bp = BeamParams(fb, d, 10**6, max(ctx.beamcap, int(table.max()), fb), 1, rerank, expand)
        _lib.check(_lib.load().ssg_learned_search_dev(
            G._handle(), self.mlp._p, self.dlandmarks.ptr, dQ, m, k, k, C.byref(bp),
            nhints, min_conf, float(quantile), table.ctypes.data, len(table), dI, dD, dcd, dcb,
            stream), "learned search")

    def confidence_dev(self, dQ, m, stream):
        """Top-landmark softmax probability of each of the m device-resident queries."""
        dH = DeviceArray((m, 1), np.uint32)
        dB = DeviceArray((m,), np.int32)
        dC = DeviceArray((m,), np.float32)
        _lib.check(_lib.load().ssg_mlp_infer_dev(self.mlp._p, dQ, m, self.dlandmarks.ptr, 1, 0.0,
                                                 dH.ptr, dB.ptr, dC.ptr, stream.handle), "infer")
        stream.synchronize()
        return dC.numpy()

    @staticmethod
    def isotonic(x, y):
        """Non-decreasing least-squares fit of y on x (pool adjacent violators): returns the
        sorted x and the fitted values."""
        o = np.argsort(x, kind="stable")
        xs, ys = np.asarray(x, np.float64)[o], np.asarray(y, np.float64)[o]
        val, wt, cnt = [], [], []
        for v in ys:
            val.append(v)
            wt.append(1.0)
            cnt.append(1)
            while len(val) > 1 and val[-2] > val[-1]:
                w = wt[-2] + wt[-1]
                val[-2] = (val[-2] * wt[-2] + val[-1] * wt[-1]) / w
                wt[-2] = w
                cnt[-2] += cnt[-1]
                val.pop()
                wt.pop()
                cnt.pop()
        return xs, np.repeat(val, cnt)

    def tune(self, ctx, Qtune, gold, target=0.95, margin=0.004, k=10, deltas=(1.05, 1.1),
             reranks=(10, 12, 14, 20), caps=(64, 128),
             quantiles=(0.5, 0.55, 0.6, 0.65, 0.7, 0.75, 0.8, 0.85, 0.9),
             fallback_ladder=(8, 12, 16, 20, 24, 28, 32, 40, 48, 64, 96, 128), verbose=None):
        """Chooses the learned knobs and the predicted-recall fallback on tuning queries.

        Qtune: (m, dim) rows; gold: (m, k) exact nearest ids (any order).  Returns a dict
        with quantile, rerank, cap, delta, min_conf (queries with a lower confidence have
        a predicted recall below `target` and fall back), fallback_bsize, the tuning recall
        and mean distance evaluations, the fallback fraction, and the calibration (confidence
        -> predicted recall).  The chosen cap is installed (self.buckets); pass min_conf and
        fallback_bsize to search_dev."""
        from . import _rows_for
        if not hasattr(self, "full_buckets"):
            self.full_buckets = self.buckets
        G = self.G
        Qr = _rows_for(G.dist, np.ascontiguousarray(Qtune, np.float32))
        m = len(Qr)
        gold = np.sort(np.asarray(gold, np.uint32)[:, :k], axis=1)
        st = Stream()
        dQ = DeviceArray.from_numpy(Qr)
        dI, dD = DeviceArray((m, k), np.uint32), DeviceArray((m, k), np.float32)
        dcd = DeviceArray((m,), np.int64)
        L = _lib.load()

        def per_query_recall(ids):
            return (np.sort(ids, axis=1)[:, :, None] == gold[:, None, :]).any(axis=2).mean(axis=1)

        conf = self.confidence_dev(dQ.ptr, m, st)
        plain_cache = {}

        def plain(rr, delta):
            if (rr, delta) not in plain_cache:
                res = None
                for b in fallback_ladder:
                    bp = BeamParams(b, delta, 10**6, max(ctx.beamcap, b), 1, rr, 1)
                    _lib.check(L.ssg_beamsearch_dev(G._handle(), dQ.ptr, m, k, k, C.byref(bp), None,
                                                    dI.ptr, dD.ptr, dcd.ptr, None, st.handle),
                               "plain search")
                    st.synchronize()
                    r, c = per_query_recall(dI.numpy()), dcd.numpy().astype(np.float64)
                    res = (b, r, c)
                    if r.mean() >= target + margin:
                        break
                plain_cache[(rr, delta)] = res
            return plain_cache[(rr, delta)]

        best = None
        for delta in deltas:
            for rr in reranks:
                fb, fr, fc = plain(rr, delta)
                for cap in caps:
                    self.buckets = tuple(min(b, cap) for b in self.full_buckets)
                    for q in quantiles:
                        self.search_dev(ctx, dQ.ptr, m, k, dI.ptr, dD.ptr, nhints=1, min_conf=-1.0,
                                        rerank=rr, quantile=q, delta=delta, fallback_bsize=fb,
                                        stream=st.handle, dcd=dcd.ptr)
                        st.synchronize()
                        r, c = per_query_recall(dI.numpy()), dcd.numpy().astype(np.float64)
                        xs, pred = self.isotonic(conf, r)
                        ok = np.nonzero(pred >= target)[0]
                        min_conf = float(xs[ok[0]]) if len(ok) else float("inf")
                        use = conf >= min_conf
                        rec = float(np.where(use, r, fr).mean())
                        cost = float(np.where(use, c, fc).mean())
                        if rec >= target + margin and (best is None or cost < best["mean_costdists"]):
                            edges = np.quantile(xs, np.linspace(0, 1, 11))
                            best = {"quantile": q, "rerank": rr, "cap": cap, "delta": delta,
                                    "min_conf": min_conf, "fallback_bsize": fb, "recall": rec,
                                    "mean_costdists": cost, "fallback_fraction": float(1 - use.mean()),
                                    "calibration": {"confidence": edges.tolist(),
                                                    "predicted_recall": np.interp(edges, xs, pred).tolist()}}
                            if verbose:
                                verbose(f"tune q={q} rr={rr} cap={cap} delta={delta}: recall {rec:.4f} "
                                        f"mean costdists {cost:.0f}, fallback {1 - use.mean():.3f} "
                                        f"(bsize {fb}, min_conf {min_conf:.3f})")
        if best is None:
            fb, fr, fc = plain(max(reranks), max(deltas))
            best = {"quantile": max(quantiles), "rerank": max(reranks), "cap": max(caps),
                    "delta": max(deltas), "min_conf": float("inf"), "fallback_bsize": fb,
                    "recall": float(fr.mean()), "mean_costdists": float(fc.mean()),
                    "fallback_fraction": 1.0, "calibration": None}
        self.buckets = tuple(min(b, best["cap"]) for b in self.full_buckets)
        self.tuned = best
        return best

    def capture(self, ctx, dQ, m, k, dI, dD, stream, **kw):
        """Records search_dev over the m device-resident queries into a HIP graph (one
        persistent search launch behind the inference, plan and ordering kernels).  The
        call is made once first so that every buffer is sized; returns a LearnedGraph."""
        self.search_dev(ctx, dQ, m, k, dI, dD, stream=stream.handle, **kw)
        stream.synchronize()
        L = _lib.load()
        _lib.check(L.ssg_capture_begin(stream.handle), "capture begin")
        try:
            self.search_dev(ctx, dQ, m, k, dI, dD, stream=stream.handle, **kw)
        finally:
            g = C.c_void_p()
            rc = L.ssg_capture_end(stream.handle, C.byref(g))
        _lib.check(rc, "capture end")
        return LearnedGraph(g)

    def search(self, ctx, Q, k, **kw):
        from . import _as_db, _qrows
        Q = _as_db(Q)
        m = len(Q)
        dQ = DeviceArray.from_numpy(_qrows(self.G, Q))
        dI = DeviceArray((m, k), np.uint32)
        dD = DeviceArray((m, k), np.float32)
        dcd = DeviceArray((m,), np.int64)
        dcb = DeviceArray((m,), np.int64)
        s = Stream()
        self.search_dev(ctx, dQ.ptr, m, k, dI.ptr, dD.ptr, stream=s.handle, dcd=dcd.ptr,
                        dcb=dcb.ptr, **kw)
        s.synchronize()
        ctx.costdists = dcd.numpy()
        ctx.costblocks = dcb.numpy()
        return dI.numpy().T, dD.numpy().T
