#!/usr/bin/env python3
"""Benchmark: QPS at recall@10 >= 0.95 of SearchGraph search on one MI355X, learned+BF16
batched form against the plain FP32 beam search.

Workload (synthetic, seeded): 1,000,000 x 128 Float32 database drawn from a
100-component isotropic Gaussian mixture (centers ~ U[-1,1]^128, sigma 0.15, equal
weights); SqL2, k = 10.  A step is one batch of 10,000 device-resident test queries from
the same mixture (`--steps` distinct batches); recall is always measured against the
exhaustive ground truth of every query.  The graph is built on the GPU by index! with the
reference's default context knobs (parallel_block = 8 x host threads = 128).

Plain form: searchbatch! with the BeamSearch that optimize_index!(MinRecall) picks on
2,000 tuning queries, raised until recall@10 >= 0.95 on 20,000 tuning queries; one
launch per step.  Its ids, distances and cost counters on the first 10,000 test queries
are compared with the oracle's restated search (`config.plain.oracle_mismatches`).

Learned+BF16 batched form: 4,096 fft landmarks; an MLP 128 -> 256 -> 256 -> (4096 + 8)
trained on the GPU on traces of 500,000 training queries; per query, the predicted
landmark as entry point and the beam size of the smallest bucket whose predicted
cumulative probability reaches a quantile; BF16 search with an exact FP32 rerank.  Its
knobs (quantile, rerank length, beam-size cap, delta) are tuned on the 20,000 tuning
queries by the product's tuner (LearnedSearch.tune), which also sets the fallback: a
query whose recall, predicted from its landmark confidence, is below the target takes the
plain BF16 search instead.  All steps run as one replay of a captured HIP graph (inference, planning,
longest-first ordering and one persistent search launch over every step's queries).
`value` is that form's QPS over the timed steps.  `config.ablation_plain_bf16` is the
same BF16 search and rerank without the network (global hints, one beam size tuned to
the same recall margin), also one graph replay: the network's own gain.

`roofline`: HBM-bound; algorithmic bytes of the learned form's search launch over its
HIP-event duration (the same launch run eagerly).  Bytes per query:
E*dim*s + B*(8 + 4*deg) + dim*4 + 8k (+ max(k, rerank)*dim*4 for the BF16 form's exact
rerank), with E = costdists, B = costblocks, s = 4 (FP32) or 2 (BF16) and deg the mean
out-degree.  `cpu_baseline`: the restated plain beam search built with the reference's
FP32 SIMD distance (oracle/, libssg_cpubase.so), OpenMP over queries, same graph and
parameters.

`sub_records` (default run; `--no-subs` skips them) hold the other configurations, each
with its own `roofline` and `cpu_baseline`: C1 exact exhaustive k-NN (100,000 x 32
U[0,1), 1,000 queries; and 1M x 128 x 10,000), C3 allknn of 1M x 96 unit vectors with
the BF16 and FP8 candidate passes, C4 200,000 mixed queries (bsize ~ U{16..256}, k ~
U{1, 10, 50}) in one graph replay on the same graph.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "similaritysearch.jl_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
MFMA_FP8_PEAK_TFLOPS = 5000.0
MFMA_FP32_PEAK_TFLOPS = 157.3
STEP = 10_000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _oracle():
    """The CPU oracle (checker) and its CPU-baseline build; loaded only by the check and
    cpu_baseline legs, never inside a timed GPU region."""
    p = os.path.join(ROOT, "oracle")
    if p not in sys.path:
        sys.path.insert(0, p)
    import pyoracle
    return pyoracle


def beam_bytes(costd, costb, deg, dim, s, k, rerank_rows=0):
    """Algorithmic HBM bytes of a beam-search launch: per query E*dim*s (rows gathered) +
    B*(8 + 4*deg) (list offsets and ids of the expanded vertices) + dim*4 (query) + 8*k
    (result) + rerank_rows*dim*4 (the BF16 form's exact rerank).  k and rerank_rows may
    be per-query arrays."""
    import numpy as np
    costd = np.asarray(costd, np.float64)
    costb = np.asarray(costb, np.float64)
    per = costd * dim * s + costb * (8.0 + 4.0 * deg) + dim * 4.0 + 8.0 * np.asarray(k, np.float64) \
        + np.asarray(rerank_rows, np.float64) * dim * 4.0
    return float(np.broadcast_to(per, costd.shape).sum())


def _traffic(name, key, applies, scale=1):
    """HBM bytes per call from a committed counter measurement (profiles/<name>) of the same
    kernel and workload, or None when it does not apply."""
    path = os.path.join(ROOT, "profiles", name)
    if not applies or not os.path.exists(path):
        return None
    with open(path) as fh:
        return int(json.load(fh)[key] * scale)


def _kname(prec, beamcap, rounds):
    """The beam kernel instantiation a launch runs (ssg_beam.h launch_v) for k <= 64: beam
    registers 1/2/4 by beam cap, gather rounds 4, 8, or 10 with one beam register in the
    BF16 form; the FP32 kernel's last parameter is its one result register."""
    nb = 1 if beamcap <= 64 else 2 if (prec and beamcap <= 128) else 4
    r = rounds if rounds == 4 or (rounds == 10 and prec and nb == 1) else 8
    return f"beam_bf16_kernel<8, {nb}, {r}>" if prec else f"beam_kernel<8, {nb}, {r}, 1>"


def _capture(L, _lib, st, fn):
    """Runs fn() once (sizing every buffer), then records it into a HIP graph."""
    fn()
    st.synchronize()
    _lib.check(L.ssg_capture_begin(st.handle), "capture begin")
    try:
        fn()
    finally:
        g = C.c_void_p()
        rc = L.ssg_capture_end(st.handle, C.byref(g))
    _lib.check(rc, "capture end")
    return g


def _time_graph(L, _lib, st, g, warmup):
    from similaritysearch.device import Event
    for _ in range(max(1, warmup)):
        _lib.check(L.ssg_graph_launch(g, st.handle), "replay")
    st.synchronize()
    e0, e1 = Event(), Event()
    t = time.perf_counter()
    e0.record(st)
    _lib.check(L.ssg_graph_launch(g, st.handle), "replay")
    e1.record(st)
    st.synchronize()
    return time.perf_counter() - t, e0.elapsed_ms(e1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--train", type=int, default=500_000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--parallel-block", type=int, default=128)
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--plain", type=str, default="", help="bsize,delta,maxvisits (skip tuning)")
    ap.add_argument("--learned", type=str, default="",
                    help="quantile,rerank,cap,delta (skip tuning)")
    ap.add_argument("--c4-queries", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-subs", action="store_true", help="headline only (no C1/C3/C4 sub-records)")
    ap.add_argument("--workload", choices=("search", "batching", "allknn-bf16", "allknn-fp8",
                                           "exhaustive", "hsp", "neardup", "copy", "gather"),
                    default="search",
                    help="search: the headline graph-search metric with sub-records (default); "
                         "batching: C4 mixed queries in one graph replay; allknn-bf16 / "
                         "allknn-fp8: exhaustive allknn with the BF16 / FP8 candidate pass and "
                         "FP32 rerank; exhaustive: exact FP32 exhaustive k-NN; hsp: hsp_queries "
                         "over a 1M allknn; neardup: neardup over 1M clustered unit vectors; copy: "
                         "1 GiB device copy; gather: rerank! over a random permutation of 1M rows "
                         "(copy and gather calibrate the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the query-sharded "
              "multi-GPU layout (one database replica per GPU) is not run here", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be >= 1")
    run = {"search": search_bench, "batching": batching_bench, "exhaustive": exhaustive_fp32,
           "hsp": hsp_bench, "neardup": neardup_bench, "copy": copy_bench,
           "gather": gather_bench,
           "allknn-bf16": lambda a: allknn_prec(a, "bf16"),
           "allknn-fp8": lambda a: allknn_prec(a, "fp8")}[args.workload]
    out = run(args)
    print(json.dumps(out), flush=True)


def search_bench(args):
    import numpy as np
    import similaritysearch as ss
    from similaritysearch import datasets, _lib
    from similaritysearch.device import DeviceArray, Event, Stream
    from similaritysearch.learned import LearnedSearch

    L = _lib.load()
    dim, k = 128, 10
    m = STEP * args.steps
    t0 = time.time()
    X = datasets.gmm(args.n, dim, seed=4)
    Q = datasets.gmm(m, dim, seed=5)
    Qtune = datasets.gmm(20_000, dim, seed=6)
    Qtrain = datasets.gmm(args.train, dim, seed=9)
    log(f"data in {time.time() - t0:.1f}s")

    t0 = time.time()
    ctx = ss.SearchGraphContext()
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), ctx,
                             parallel_block=args.parallel_block, seed=1, verbose=25)
    build_s = time.time() - t0
    log(f"index! {build_s:.1f}s nnz={len(G.adj.end_point)} hints={len(G.hints)} {G.algo}")
    h = G._handle()
    deg = len(G.adj.end_point) / args.n

    def exact(Qr):
        gi = np.zeros((len(Qr), k), np.uint32)
        gd = np.zeros((len(Qr), k), np.float32)
        _lib.check(L.ssg_exhaustive_knn(h, np.ascontiguousarray(Qr).ctypes.data, len(Qr), k,
                                        gi.ctypes.data, gd.ctypes.data), "exhaustive")
        return np.sort(gi, axis=1)
    gold, gold_tune = exact(Q), exact(Qtune)

    def recall(g, ids):
        """macrorecall (src/perf.jl) for rows of ids against rows of sorted gold ids"""
        return float((np.sort(ids, axis=1)[:, :, None] == g[:, None, :]).any(axis=2).sum() / g.size)

    st = Stream()
    sp = st.handle
    dQ = DeviceArray.from_numpy(Q)
    dQt = DeviceArray.from_numpy(Qtune)
    dI = DeviceArray((m, k), np.uint32)
    dD = DeviceArray((m, k), np.float32)
    dcd = DeviceArray((m,), np.int64)
    dcb = DeviceArray((m,), np.int64)

    def plain_run(bp, dq, mm, off=0):
        _lib.check(L.ssg_beamsearch_dev(h, dq.ptr + off * dim * 4, mm, k, k, C.byref(bp), None,
                                        dI.ptr + off * k * 4, dD.ptr + off * k * 4,
                                        dcd.ptr + off * 8, dcb.ptr + off * 8, sp), "searchbatch!")

    # ------------------------------------------------------------------ plain form
    def plain_recall(algo, c=ctx):
        G.algo = algo
        plain_run(G.beam_params(c), dQt, len(Qtune))
        st.synchronize()
        return recall(gold_tune, dI.numpy()[:len(Qtune)]), float(dcd.numpy()[:len(Qtune)].mean())

    if args.plain:
        b, d, mv = args.plain.split(",")
        palgo = ss.BeamSearch(int(b), float(d), int(mv))
    else:
        palgo, _ = G.optimize(ctx, 0.9, ss.MatrixDatabase.from_rows(Qtune[:2000]))
        prec, _ = plain_recall(palgo)
        log(f"MinRecall(0.9): {palgo} tuning recall@10={prec:.4f}")
        target = args.target_recall
        while prec < args.target_recall and target < 1.0:
            palgo, _ = G.optimize(ctx, min(target, 0.999), ss.MatrixDatabase.from_rows(Qtune[:2000]))
            prec, _ = plain_recall(palgo)
            log(f"MinRecall({target:.3f}): {palgo} tuning recall@10={prec:.4f}")
            target += 0.01
        while prec < args.target_recall and palgo.bsize < 256:
            palgo = ss.BeamSearch(palgo.bsize * 2, palgo.delta, 10**6)
            prec, _ = plain_recall(palgo)
            log(f"widened: {palgo} tuning recall@10={prec:.4f}")
    G.algo = palgo
    bp_plain = G.beam_params(ctx)
    # launch shape from the product's autotuner (ssg_optimize_launch) on the tuning queries
    pshape, _, _ = G.optimize_launch_dev(bp_plain, dQt.ptr, len(Qtune), k, k, gold_tune,
                                         args.target_recall)
    log(f"plain launch shape (slots per CU, gather rounds): {pshape}")
    for w in range(args.warmup):
        plain_run(bp_plain, dQ, STEP, (w % args.steps) * STEP)
    st.synchronize()
    t = time.perf_counter()
    for s_ in range(args.steps):
        plain_run(bp_plain, dQ, STEP, s_ * STEP)
    st.synchronize()
    pwall = time.perf_counter() - t
    plain_qps = m / pwall
    pI, pD = dI.numpy().copy(), dD.numpy().copy()
    prec_dev = recall(gold, pI)
    pcostd, pcostb = dcd.numpy().copy(), dcb.numpy().copy()
    log(f"plain: {plain_qps:.0f} QPS, recall@10 {prec_dev:.4f}")

    # ------------------------------------------------------------------ learned form
    t0 = time.time()
    LS = LearnedSearch(G, nlandmarks=4096, hidden=256)
    LS.landmark_labels()
    losses = LS.fit(Qtrain, ctx, epochs=args.epochs, batch=2048, verbose=log)
    fit_s = time.time() - t0
    log(f"learned model in {fit_s:.1f}s, losses {np.round(losses, 3).tolist()}")
    full_buckets = LS.buckets
    margin = 0.004
    if args.learned:
        q, rr, cap, delta = args.learned.split(",")
        q, rr, cap, delta = float(q), int(rr), int(cap), float(delta)
        LS.buckets = tuple(min(b, cap) for b in full_buckets)
        tuned = {"quantile": q, "rerank": rr, "cap": cap, "delta": delta, "min_conf": -1.0,
                 "fallback_bsize": palgo.bsize, "fallback_fraction": None, "calibration": None}
    else:
        # the product's tuner: cheapest setting whose tuning recall clears the target with a
        # margin, queries of predicted recall below the target falling back to plain search
        t0 = time.time()
        tuned = LS.tune(ctx, Qtune, gold_tune, target=args.target_recall, margin=margin, k=k,
                        verbose=log)
        log(f"tuned in {time.time() - t0:.1f}s: {tuned}")
    q, rr, cap, delta = tuned["quantile"], tuned["rerank"], tuned["cap"], tuned["delta"]
    min_conf, fb = tuned["min_conf"], tuned["fallback_bsize"]
    kw = dict(nhints=1, rerank=rr, quantile=q, delta=delta, min_conf=min_conf, fallback_bsize=fb)
    log(f"learned: quantile={q} rerank={rr} cap={cap} delta={delta} min_conf={min_conf} fallback bsize {fb}")

    # launch shape for the learned search, from the autotuner over the tuning queries with
    # their planned hints and beam sizes (the rerank and wide-step knobs it may also raise)
    mt = len(Qtune)
    tH, tB, tC = DeviceArray((mt, 1), np.uint32), DeviceArray((mt,), np.int32), DeviceArray((mt,), np.float32)
    tN, tS = DeviceArray((mt,), np.int32), DeviceArray((mt,), np.int32)
    ttab = LS.table(0)
    _lib.check(L.ssg_mlp_infer_dev(LS.mlp._p, dQt.ptr, mt, LS.dlandmarks.ptr, 1, float(q), tH.ptr,
                                   tB.ptr, tC.ptr, sp), "infer")
    _lib.check(L.ssg_learned_plan_dev(h, tB.ptr, tC.ptr, mt, 1, float(min_conf), ttab.ctypes.data,
                                      len(ttab), fb, tN.ptr, tS.ptr, sp), "plan")
    st.synchronize()
    bp_t = ss.BeamParams(fb, delta, 10**6, max(ctx.beamcap, int(ttab.max()), fb), 1, rr, 1)
    lshape, lrec_t, _ = G.optimize_launch_dev(bp_t, dQt.ptr, mt, k, k, gold_tune,
                                              args.target_recall + margin,
                                              ss.QueryOpts(None, tS.ptr, tH.ptr, tN.ptr, 1, None))
    if bp_t.rerank != rr or bp_t.expand != 1:
        log(f"launch tuner raised rerank {rr} -> {bp_t.rerank}, expand {bp_t.expand}")
    rr, expand = bp_t.rerank, bp_t.expand
    kw.update(rerank=rr, expand=expand)
    log(f"learned launch shape {lshape}, tuning recall {lrec_t:.4f}")
    del tH, tB, tC, tN, tS

    # batched: one graph replay covers every step
    g = LS.capture(ctx, dQ.ptr, m, k, dI.ptr, dD.ptr, st, **kw)
    for _ in range(max(1, args.warmup)):
        g.launch(st)
    st.synchronize()
    e0, e1 = Event(), Event()
    t = time.perf_counter()
    e0.record(st)
    g.launch(st)
    e1.record(st)
    st.synchronize()
    lwall = time.perf_counter() - t
    lms = e0.elapsed_ms(e1)
    learned_qps = m / lwall
    lrec = recall(gold, dI.numpy())
    g.close()
    log(f"learned batched: {learned_qps:.0f} QPS ({lms:.3f} ms by events), recall@10 {lrec:.4f}")

    # the same pipeline eagerly, with the search launch timed alone for the roofline
    nb = len(LS.buckets)
    dH = DeviceArray((m, 1), np.uint32)
    dB = DeviceArray((m,), np.int32)
    dC = DeviceArray((m,), np.float32)
    dN = DeviceArray((m,), np.int32)
    dS = DeviceArray((m,), np.int32)
    dO = DeviceArray((m,), np.uint32)
    table = LS.table(0)
    bp_l = ss.BeamParams(fb, delta, 10**6, max(ctx.beamcap, int(table.max()), fb), 1, rr, expand)
    qo = ss.QueryOpts(None, dS.ptr, dH.ptr, dN.ptr, 1, dO.ptr)
    e2, e3, e4 = Event(), Event(), Event()
    e2.record(st)
    _lib.check(L.ssg_mlp_infer_dev(LS.mlp._p, dQ.ptr, m, LS.dlandmarks.ptr, 1, float(q), dH.ptr,
                                   dB.ptr, dC.ptr, sp), "infer")
    _lib.check(L.ssg_learned_plan_dev(h, dB.ptr, dC.ptr, m, 1, float(min_conf), table.ctypes.data, nb,
                                      fb, dN.ptr, dS.ptr, sp), "plan")
    _lib.check(L.ssg_order_by_bsize_dev(h, dS.ptr, m, dO.ptr, sp), "order")
    e3.record(st)
    _lib.check(L.ssg_beamsearch_dev(h, dQ.ptr, m, k, k, C.byref(bp_l), C.byref(qo), dI.ptr,
                                    dD.ptr, dcd.ptr, dcb.ptr, sp), "learned search")
    e4.record(st)
    st.synchronize()
    pre_ms, sms = e2.elapsed_ms(e3), e3.elapsed_ms(e4)
    lcostd, lcostb = dcd.numpy().copy(), dcb.numpy().copy()
    fb_frac = float((dN.numpy() < 0).mean())

    # per-step (unbatched) learned launches, for reference
    for w in range(args.warmup):
        LS.search_dev(ctx, dQ.ptr + (w % args.steps) * STEP * dim * 4, STEP, k, dI.ptr, dD.ptr,
                      stream=sp, **kw)
    st.synchronize()
    t = time.perf_counter()
    for s_ in range(args.steps):
        LS.search_dev(ctx, dQ.ptr + s_ * STEP * dim * 4, STEP, k, dI.ptr + s_ * STEP * k * 4,
                      dD.ptr + s_ * STEP * k * 4, stream=sp, **kw)
    st.synchronize()
    step_qps = m / (time.perf_counter() - t)

    # ------------------------------------------------------------------ ablation: no network
    # the same BF16 search and rerank from the global hints with one beam size, the smallest
    # reaching the learned form's tuning margin (the tuner's fallback search) for every
    # query; all steps in one graph replay
    ab = fb
    actx = ss.SearchGraphContext(beamcap=max(32, ab), precision="bf16", rerank=rr)
    G.algo = ss.BeamSearch(ab, delta, 10**6)
    bp_a = G.beam_params(actx)
    ashape, _, _ = G.optimize_launch_dev(bp_a, dQt.ptr, len(Qtune), k, k, gold_tune,
                                         args.target_recall + margin)
    ga = _capture(L, _lib, st, lambda: plain_run(bp_a, dQ, m))
    awall, ams = _time_graph(L, _lib, st, ga, args.warmup)
    L.ssg_graph_destroy(ga)
    arec = recall(gold, dI.numpy())
    acd = float(dcd.numpy().mean())
    ablation = {"qps": round(m / awall, 1), "recall@10": round(arec, 4), "bsize": ab, "delta": delta,
                "rerank": bp_a.rerank, "expand": bp_a.expand, "launch_shape": list(ashape),
                "mean_costdists": round(acd, 1), "graph_replay_ms": round(ams, 3),
                "learned_over_this": round(learned_qps / (m / awall), 3),
                "costdists_saved_by_network": round(1.0 - float(lcostd.mean()) / acd, 4)}
    log(f"ablation plain BF16 batched: {m / awall:.0f} QPS, recall@10 {arec:.4f}, bsize {ab}")
    G.algo = palgo

    # HBM traffic of the learned search launch from the committed counter measurement of
    # the same kernel and workload (FETCH_SIZE / WRITE_SIZE passes scaled by a 1 GiB copy)
    traffic = _traffic("beam_search_learned_bf16_1m128_traffic.json", "hbm_bytes_per_query",
                       args.n == 1_000_000, scale=m)
    lbytes = beam_bytes(lcostd, lcostb, deg, dim, 2, k, max(k, rr))
    achieved = lbytes / (sms * 1e-3) / 1e9
    pbytes = beam_bytes(pcostd, pcostb, deg, dim, 4, k)

    # ------------------------------------------------------------------ oracle check, CPU baseline
    po = _oracle()
    threads = _threads()
    offs = np.concatenate([[0], G.adj.offset]).astype(np.int64)
    Qc = Q[:STEP]
    oi, od, ocd, ocb = po.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Qc, k, palgo.bsize,
                                     palgo.delta, palgo.maxvisits, nthreads=threads)
    mism = int(((oi != pI[:STEP]).any(1) | (od != pD[:STEP]).any(1) | (ocd != pcostd[:STEP])
                | (ocb != pcostb[:STEP])).sum())
    orec, grec = recall(gold[:STEP], oi), recall(gold[:STEP], pI[:STEP])
    log(f"oracle check on {STEP} plain queries: {mism} mismatching queries; recall oracle {orec:.4f} "
        f"GPU {grec:.4f}")
    cpu = None
    if not args.no_cpu:
        po.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Qc[:200], k, palgo.bsize,
                      palgo.delta, palgo.maxvisits, nthreads=threads, baseline=True)
        t0 = time.perf_counter()
        reps = 0
        while True:
            ci, _, _, _ = po.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Qc, k,
                                        palgo.bsize, palgo.delta, palgo.maxvisits,
                                        nthreads=threads, baseline=True)
            reps += 1
            if time.perf_counter() - t0 > 10.0 or reps >= 20:
                break
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(STEP * reps / cpu_s, 1), "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": f"restated plain searchbatch with the reference's FP32 SIMD distance "
                         f"(vecs.jl @fastmath @simd) over the first {STEP} test queries x{reps}, "
                         f"same graph and BeamSearch as the plain GPU form, OpenMP over queries; "
                         f"recall@10 {recall(gold[:STEP], ci):.4f}"}

    out = {
        "metric": "QPS at recall@10>=0.95, 1M×128 SqL2, 1 GPU: learned+BF16 batched vs plain beam search",
        "value": round(learned_qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(lwall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic: seeded 100-component Gaussian mixture, sigma 0.15; 10,000 distinct test "
                "queries per step; 500,000 training and 20,000 tuning queries for the learned form",
        "config": {
            "workload": "SearchGraph search, 1,000,000x128 GMM database, 10,000 queries per step, "
                        "SqL2 (squared L2: Δ acts on squared distances), k=10, recall@10 >= 0.95 "
                        "against exhaustive ground truth",
            "form": "learned entry points and beam sizes + BF16 beam search + FP32 rerank, all steps "
                    "in one HIP graph replay (one persistent search launch)",
            "recall@10": round(lrec, 4),
            "learned": {"quantile": q, "rerank": rr, "expand": expand, "beam_cap": cap,
                        "delta": delta, "nhints": 1, "launch_shape": list(lshape),
                        "min_conf": min_conf, "fallback_bsize": fb,
                        "fallback_fraction_test": round(fb_frac, 4),
                        "fallback_fraction_tune": tuned["fallback_fraction"],
                        "predicted_recall_calibration": tuned["calibration"],
                        "buckets": list(full_buckets), "landmarks": LS.L,
                        "mean_costdists": round(float(lcostd.mean()), 1),
                        "mean_costblocks": round(float(lcostb.mean()), 1),
                        "graph_replay_ms": round(lms, 3), "infer_plan_order_ms": round(pre_ms, 3),
                        "search_ms": round(sms, 3), "per_step_launch_qps": round(step_qps, 1),
                        "train_queries": args.train, "epochs": args.epochs, "fit_s": round(fit_s, 1)},
            "plain": {"qps": round(plain_qps, 1), "recall@10": round(prec_dev, 4),
                      "launch_shape": list(pshape),
                      "bsize": palgo.bsize, "delta": palgo.delta, "maxvisits": palgo.maxvisits,
                      "precision": "fp32", "launch": "one per step",
                      "mean_costdists": round(float(pcostd.mean()), 1),
                      "mean_costblocks": round(float(pcostb.mean()), 1),
                      "algorithmic_GB_per_step": round(pbytes / args.steps / 1e9, 3),
                      "oracle_checked_queries": STEP, "oracle_mismatches": mism,
                      "recall@10_oracle": round(orec, 4), "recall@10_gpu_same_queries": round(grec, 4),
                      "recall_delta": round(grec - orec, 5)},
            "ablation_plain_bf16": ablation,
            "learned_over_plain": round(learned_qps / plain_qps, 3),
            "n": args.n, "queries": m, "k": k, "beamcap": ctx.beamcap,
            "parallel_block": args.parallel_block, "graph_nnz": int(len(G.adj.end_point)),
            "hints": int(len(G.hints)), "build_s": round(build_s, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/beam_search_learned_bf16_1m128_traffic.md" if traffic else None,
                     "kernel": _kname(1, bp_l.beamcap, lshape[1]),
                     "kernel_ms": round(sms, 3), "algorithmic_bytes_per_launch": int(lbytes)},
        "cpu_baseline": cpu,
    }
    if not args.no_subs:
        del LS, dH, dB, dC, dN, dS, dO
        subs = {}
        log("sub-record C4 (batching)")
        subs["c4_batching"] = batching_run(args, G, X, compact=True)
        G.close()
        log("sub-record C1 (exhaustive)")
        subs["c1_exhaustive"] = exhaustive_fp32(args, compact=True)
        X96 = datasets.unit(args.n, 96, seed=6)
        for fmt in ("bf16", "fp8"):
            log(f"sub-record C3 (allknn {fmt})")
            subs[f"c3_allknn_{fmt}"] = allknn_prec(args, fmt, X96, compact=True)
        out["sub_records"] = subs
    return out


def batching_bench(args):
    """C4 alone: builds the 1M x 128 graph (as the headline does), then batching_run."""
    import similaritysearch as ss
    from similaritysearch import datasets
    X = datasets.gmm(args.n, 128, seed=4)
    t0 = time.time()
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), ss.SearchGraphContext(),
                             parallel_block=args.parallel_block, seed=1)
    log(f"index! {time.time() - t0:.1f}s")
    return batching_run(args, G, X, compact=False)


def batching_run(args, G, X, compact):
    """C4, the batching form: 200,000 queries from the mixture (seed 7) with per-query beam
    sizes ~ U{16, 32, 64, 128, 256} and k ~ U{1, 10, 50} (seed 8) against the 1M x 128
    graph, all in one persistent launch captured in a HIP graph (longest-first order).
    FP32 (`value`) and the BF16 form (rerank 20 or more).  The launch shape (wave slots per
    CU, gather rounds) and the BF16 rerank / wide steps come from the product's autotuner
    (ssg_optimize_launch) on 20,000 tuning queries drawn the same way (seeds 17, 18) with a
    0.95 recall target.  Recall: per query the share of its exact
    min(k, 10) nearest neighbours among its first min(k, 10) results, averaged.  Checked
    against the oracle on 2,000 queries; `cpu_baseline`: the restated search with the
    reference's FP32 SIMD distance, per-query k and bsize, 256-row beam buffers."""
    import numpy as np
    import similaritysearch as ss
    from similaritysearch import datasets, _lib
    from similaritysearch.device import DeviceArray, Stream

    L = _lib.load()
    h = G._handle()
    dim, KS = 128, 50
    m = args.c4_queries
    n = len(X)
    deg = len(G.adj.end_point) / n
    Q = datasets.gmm(m, dim, seed=7)
    rng = np.random.default_rng(8)
    bq = rng.choice([16, 32, 64, 128, 256], m).astype(np.int32)
    kq = rng.choice([1, 10, 50], m).astype(np.int32)
    gold = np.zeros((m, 10), np.uint32)
    gd = np.zeros((m, 10), np.float32)
    for s in range(0, m, 50_000):
        e = min(m, s + 50_000)
        _lib.check(L.ssg_exhaustive_knn(h, np.ascontiguousarray(Q[s:e]).ctypes.data, e - s, 10,
                                        gold[s:e].ctypes.data, gd[s:e].ctypes.data), "exhaustive")
    kk = np.minimum(kq, 10)

    def recall(ids):
        hit = 0.0
        for c in (1, 10):
            sel = kk == c
            a, b = ids[sel, :c], gold[sel, :c]
            hit += float((a[:, :, None] == b[:, None, :]).any(2).sum(1).astype(np.float64).sum() / c)
        return hit / m

    st = Stream()
    dQ, dk, db = DeviceArray.from_numpy(Q), DeviceArray.from_numpy(kq), DeviceArray.from_numpy(bq)
    dI, dD = DeviceArray((m, KS), np.uint32), DeviceArray((m, KS), np.float32)
    dcd, dcb, dO = DeviceArray((m,), np.int64), DeviceArray((m,), np.int64), DeviceArray((m,), np.uint32)
    _lib.check(L.ssg_order_by_bsize_dev(h, db.ptr, m, dO.ptr, st.handle), "order")
    qo = ss.QueryOpts(dk.ptr, db.ptr, None, None, 0, dO.ptr)
    # tuning queries drawn the same way (seeds 17, 18) for the launch autotuner
    mt = 20_000
    Qt = datasets.gmm(mt, dim, seed=17)
    trng = np.random.default_rng(18)
    tbq = trng.choice([16, 32, 64, 128, 256], mt).astype(np.int32)
    tkq = trng.choice([1, 10, 50], mt).astype(np.int32)
    tgold = np.zeros((mt, 10), np.uint32)
    tgd = np.zeros((mt, 10), np.float32)
    _lib.check(L.ssg_exhaustive_knn(h, np.ascontiguousarray(Qt).ctypes.data, mt, 10, tgold.ctypes.data,
                                    tgd.ctypes.data), "exhaustive")
    dQt, dtk, dtb = DeviceArray.from_numpy(Qt), DeviceArray.from_numpy(tkq), DeviceArray.from_numpy(tbq)
    qot = ss.QueryOpts(dtk.ptr, dtb.ptr, None, None, 0, None)
    forms = {}
    for name, prec, rr in (("fp32", 0, 0), ("bf16", 1, 20)):
        bp = ss.BeamParams(16, G.algo.delta, 10**6, 256, prec, rr, 1)
        shape, trec, _ = G.optimize_launch_dev(bp, dQt.ptr, mt, KS, KS, tgold, 0.95, qot)
        log(f"C4 {name}: launch shape {shape}, rerank {bp.rerank}, expand {bp.expand}, "
            f"tuning recall {trec:.4f}")

        def run():
            _lib.check(L.ssg_beamsearch_dev(h, dQ.ptr, m, KS, KS, C.byref(bp), C.byref(qo), dI.ptr,
                                            dD.ptr, dcd.ptr, dcb.ptr, st.handle), "searchbatch!")
        g = _capture(L, _lib, st, run)
        wall, ms = _time_graph(L, _lib, st, g, 1 if compact else args.warmup)
        L.ssg_graph_destroy(g)
        ids, cd, cb = dI.numpy(), dcd.numpy(), dcb.numpy()
        nbytes = beam_bytes(cd, cb, deg, dim, 4 if prec == 0 else 2, kq,
                            np.maximum(kq, bp.rerank) if prec else 0)
        forms[name] = {"qps": round(m / wall, 1), "graph_replay_ms": round(ms, 3),
                       "recall": round(recall(ids), 4), "launch_shape": list(shape),
                       "rerank": bp.rerank, "expand": bp.expand,
                       "recall@10_k>=10": round(float((ids[kq >= 10, :10][:, :, None] ==
                                                       gold[kq >= 10][:, None, :]).any(2).mean()), 4),
                       "mean_costdists": round(float(cd.mean()), 1),
                       "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                       "algorithmic_bytes_per_launch": int(nbytes)}
        if prec == 0:
            I32, D32, CD32, CB32 = ids.copy(), dD.numpy().copy(), cd.copy(), cb.copy()
        log(f"C4 {name}: {m / wall:.0f} QPS ({ms:.2f} ms), recall {forms[name]['recall']:.4f}")

    po = _oracle()
    threads = _threads()
    offs = np.concatenate([[0], G.adj.offset]).astype(np.int64)
    nc = 2000
    oi, od, ocd, ocb = po.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q[:nc], 10, 16,
                                     G.algo.delta, 10**6, beamcap=256, kq=kq[:nc], bq=bq[:nc],
                                     nthreads=threads)   # the FP32 form (bit-exact at any shape)
    ks = oi.shape[1]
    mism = int(((oi != I32[:nc, :ks]).any(1) | (od != D32[:nc, :ks]).any(1) | (ocd != CD32[:nc])
                | (ocb != CB32[:nc])).sum())
    log(f"C4 oracle check on {nc} queries: {mism} mismatching")
    cpu = None
    if not args.no_cpu:
        po.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q[:100], 10, 16, G.algo.delta,
                      10**6, beamcap=256, kq=kq[:100], bq=bq[:100], nthreads=threads, baseline=True)
        nq, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < (8.0 if compact else 15.0) and nq < m:
            e = min(m, nq + 2000)
            po.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q[nq:e], 10, 16, G.algo.delta,
                          10**6, beamcap=256, kq=kq[nq:e], bq=bq[nq:e], nthreads=threads,
                          baseline=True)
            nq = e
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(nq / cpu_s, 1), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"restated searchbatch! with per-query KnnSorted k and bsize (256-row beam "
                         f"buffers) and the reference's FP32 SIMD distance over the first {nq} of the "
                         f"{m} queries, OpenMP over queries"}
    f = forms["fp32"]
    return {
        "metric": "QPS of 200,000 mixed queries (bsize ~U{16..256}, k ~U{1,10,50}) in one graph "
                  "replay, 1M×128 SqL2",
        "value": f["qps"], "unit": "queries/s", "ms_per_step": f["graph_replay_ms"], "steps": 1,
        "dtype": "f32",
        "config": {"workload": f"searchbatch! queue-vector form, {m} queries from the 1Mx128 mixture, "
                               f"per-query bsize and k, longest-first order, one persistent launch",
                   "forms": forms, "oracle_checked_queries": nc, "oracle_mismatches": mism,
                   "delta": G.algo.delta, "beamcap": 256},
        "roofline": {"bound": "hbm", "achieved": f["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(f["achieved_GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": _traffic("beam_search_batching_1m128_traffic.json", "hbm_bytes_per_launch",
                                         n == 1_000_000 and m == 200_000),
                     "kernel": _kname(0, 256, f["launch_shape"][1]), "kernel_ms": f["graph_replay_ms"],
                     "algorithmic_bytes_per_launch": f["algorithmic_bytes_per_launch"]},
        "cpu_baseline": cpu,
    }


def exhaustive_fp32(args, compact=False):
    """Exact exhaustive k-NN (ExhaustiveSearch + searchbatch, k = 10, SqL2): the 100,000 x 32
    U[0,1) database (seed 1) with 1,000 queries (seed 2), the reference's CPU-runnable
    configuration, and the 1,000,000 x 128 GMM database (seed 4) with 10,000 queries (seed 5)
    per step.  `value` is queries/s on the first in the compact sub-record and on the second
    as the standalone workload; each shape carries its own `roofline` (2 * dim * n * m MFMA
    flops over the whole call against the FP32 dense peak).  Results are checked against
    the oracle on 200 queries of each; `cpu_baseline`: the restated ExhaustiveSearch with the
    reference's FP32 SIMD distance, OpenMP over queries, on the 100,000 x 32 configuration."""
    import numpy as np
    import similaritysearch as ss
    from similaritysearch import datasets, _lib
    from similaritysearch.device import DeviceArray, Event, Stream

    L = _lib.load()
    k = 10
    steps = min(args.steps, 5) if compact else args.steps
    st = Stream()
    sp = st.handle
    po = _oracle()
    threads = _threads()
    res = {}
    for name, n, m, dim, seeds in [("uniform_100k32", 100_000, 1_000, 32, (1, 2)),
                                   ("gmm_1m128", args.n, 10_000, 128, (4, 5))]:
        if name == "uniform_100k32":
            X, Q = datasets.uniform(n, dim, seeds[0]), datasets.uniform(m, dim, seeds[1])
        else:
            X, Q = datasets.gmm(n, dim, seed=seeds[0]), datasets.gmm(m, dim, seed=seeds[1])
        E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
        h = E._handle()
        dQ = DeviceArray.from_numpy(Q)
        dI = DeviceArray((m, k), np.uint32)
        dD = DeviceArray((m, k), np.float32)

        def run():
            _lib.check(L.ssg_exhaustive_knn_dev(h, dQ.ptr, m, k, dI.ptr, dD.ptr, sp), "searchbatch!")
        for _ in range(max(1, args.warmup)):
            run()
        st.synchronize()
        e0, e1 = Event(), Event()
        e0.record(st)
        for _ in range(steps):
            run()
        e1.record(st)
        st.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        gi, gd = po.exhaustive("SqL2", X, Q[:200], k, nthreads=threads)
        ok = bool((dI.numpy()[:200] == gi).all() and (dD.numpy()[:200] == gd).all())
        tf = 2.0 * dim * n * m / (ms * 1e-3) / 1e12
        log(f"exhaustive {name}: {ms:.3f} ms per {m} queries, {tf:.1f} TFLOP/s, oracle match on 200 queries: {ok}")
        res[name] = {"ms": round(ms, 3), "qps": round(m / (ms * 1e-3), 1), "n": n, "m": m, "dim": dim,
                     "oracle_match_200": ok,
                     "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": MFMA_FP32_PEAK_TFLOPS,
                                  "unit": "TFLOP/s", "frac": round(tf / MFMA_FP32_PEAK_TFLOPS, 4),
                                  "traffic": (_traffic("exhaustive_fp32_1m128_traffic.json",
                                                       "hbm_bytes_per_call",
                                                       name == "gmm_1m128" and n == 1_000_000)
                                              or _traffic("exhaustive_fp32_100k32_traffic.json",
                                                          "hbm_bytes_per_call",
                                                          name == "uniform_100k32" and n == 100_000)),
                                  "kernel": "exh_stage1_l<%d> + exh_stage2 (whole call)" % ((dim + 15) // 16 * 16),
                                  "kernel_ms": round(ms, 3), "flops_per_launch": 2.0 * dim * n * m}}
        if name == "uniform_100k32" and not args.no_cpu:
            po.exhaustive("SqL2", X, Q[:64], k, nthreads=threads, baseline=True)
            reps, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < (5.0 if compact else 15.0) and reps < 50:
                po.exhaustive("SqL2", X, Q, k, nthreads=threads, baseline=True)
                reps += 1
            cpu_s = time.perf_counter() - t0
            res["cpu"] = {"value": round(m * reps / cpu_s, 1), "unit": "queries/s", "cores": threads,
                          "kind": "port",
                          "sample": f"restated ExhaustiveSearch (ascending scan, KnnSorted) with the "
                                    f"reference's FP32 SIMD distance, all {m} queries against the "
                                    f"100,000 x 32 database x{reps}, OpenMP over queries"}
        E.close()
    first = "uniform_100k32" if compact else "gmm_1m128"
    main = res[first]
    return {
        "metric": f"exact exhaustive k-NN queries/s, {main['n']}×{main['dim']} SqL2 k=10 "
                  f"({main['m']} queries per step)",
        "value": main["qps"], "unit": "queries/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": main["ms"], "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: seeded U[0,1) (100k x 32) and GMM (1M x 128) rows",
        "config": {"workload": "ExhaustiveSearch searchbatch, exact FP32 canonical distances, k=10",
                   "uniform_100k32": res["uniform_100k32"], "gmm_1m128": res["gmm_1m128"]},
        "roofline": main["roofline"],
        "cpu_baseline": res.get("cpu"),
    }


def allknn_prec(args, fmt, X=None, compact=False):
    """allknn (k = 32) of 1,000,000 unit vectors (96 dims, N(0, I) seed 6, L2-normalised),
    NormCosine, with the precision form: BF16 or FP8 (fmt) candidate pass on the matrix
    cores keeping 128 candidates per object, then rerank! with the FP32 canonical
    distance.  A step is one whole allknn; `value` is objects (queries) per second.
    `roofline`: the candidate pass is MFMA-bound at 2 * 96 * n^2 flops per allknn (the
    FP8 pass pads rows to 128 values; the padding is not counted), priced against the
    BF16 or FP8 dense peak over the time of the whole call (rerank included, so the
    fraction is a floor).  Recall and errors are measured on 10,000 evenly spaced objects
    against the exact FP32 exhaustive search (recall) and FP64 (errors).  `cpu_baseline`:
    the restated ExhaustiveSearch with the reference's FP32 SIMD distance on a bounded
    sample of the objects."""
    import numpy as np
    import similaritysearch as ss
    from similaritysearch import datasets, _lib
    from similaritysearch.device import DeviceArray, Event, Stream

    L = _lib.load()
    n = args.n
    dim, k, ncand = 96, 32, 128
    steps = min(args.steps, 3) if compact else args.steps
    if X is None:
        t0 = time.time()
        X = datasets.unit(n, dim, seed=6)
        log(f"data in {time.time() - t0:.1f}s")
    E = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
    h = E._handle()
    st = Stream()
    sp = st.handle
    dI = DeviceArray((n, k), np.uint32)
    dD = DeviceArray((n, k), np.float32)
    dCI = DeviceArray((n, ncand), np.uint32)
    dCD = DeviceArray((n, ncand), np.float32)

    pcode = {"bf16": 1, "fp8": 2}[fmt]  # SSG_BF16, SSG_FP8

    def run(cands):
        _lib.check(L.ssg_allknn_exhaustive_prec_dev(h, k, pcode, ncand, dI.ptr, dD.ptr,
                                                     dCI.ptr if cands else None,
                                                     dCD.ptr if cands else None, sp), "allknn")

    for _ in range(1 if compact else max(1, args.warmup)):
        run(False)
    st.synchronize()
    e0, e1 = Event(), Event()
    e0.record(st)
    for _ in range(steps):
        run(False)
    e1.record(st)
    st.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    log(f"allknn-{fmt}: {ms:.2f} ms per allknn of {n}")
    run(True)
    st.synchronize()
    ids, dists = dI.numpy(), dD.numpy()
    cids, cdists = dCI.numpy(), dCD.numpy()

    # recall and errors on 10,000 evenly spaced objects
    sub = np.linspace(0, n - 1, min(n, 10_000)).astype(np.int64)
    gi = np.zeros((len(sub), k), np.uint32)
    gd = np.zeros((len(sub), k), np.float32)
    Qs = np.ascontiguousarray(X[sub])
    _lib.check(L.ssg_exhaustive_knn(h, Qs.ctypes.data, len(sub), k, gi.ctypes.data, gd.ctypes.data),
               "exhaustive")
    rec = float((np.sort(ids[sub], 1)[:, :, None] == np.sort(gi, 1)[:, None, :]).any(2).sum() / gi.size)
    X64 = X.astype(np.float64) if not compact else None
    Xs64 = X[sub].astype(np.float64)

    def d64(idm):
        rows = X[idm.astype(np.int64) - 1].astype(np.float64) if X64 is None else X64[idm.astype(np.int64) - 1]
        return 1.0 - np.einsum("qd,qkd->qk", Xs64, rows)
    err_final = np.abs(dists[sub].astype(np.float64) - d64(ids[sub]))
    err_cand = np.abs(cdists[sub].astype(np.float64) - d64(cids[sub]))
    log(f"recall@{k} {rec:.5f}; final |err| max {err_final.max():.3e} mean {err_final.mean():.3e}; "
        f"candidate |err| max {err_cand.max():.3e} mean {err_cand.mean():.3e}")
    E.close()
    del dI, dD, dCI, dCD

    flops = 2.0 * dim * float(n) * float(n)
    tflops = flops / (ms * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        po = _oracle()
        threads = _threads()
        po.exhaustive("NormCosine", X, Qs[:16], k, nthreads=threads, baseline=True)
        nq, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < (6.0 if compact else 15.0) and nq < len(sub):
            po.exhaustive("NormCosine", X, Qs[nq:nq + 128], k, nthreads=threads, baseline=True)
            nq += 128
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(nq / cpu_s, 2), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"restated ExhaustiveSearch (ascending scan, KnnSorted) with the reference's "
                         f"FP32 SIMD distance, {nq} of the objects against all {n}, OpenMP over "
                         f"queries; a whole allknn would take {n / (nq / cpu_s) / 60:.0f} min"}

    FMT = fmt.upper()
    peak = MFMA_BF16_PEAK_TFLOPS if fmt == "bf16" else MFMA_FP8_PEAK_TFLOPS
    return {
        "metric": f"allknn objects/s, 1M×96 NormCosine k=32, {FMT} candidate pass (128 per object) + FP32 rerank",
        "value": round(n / (ms * 1e-3), 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": fmt,
        "data": "synthetic: N(0, I) rows, seed 6, L2-normalised",
        "config": {"workload": f"allknn over ExhaustiveSearch, {n}x{dim} NormCosine, k={k}, "
                               f"{ncand} {FMT} candidates per object reranked in FP32",
                   f"recall@{k}_vs_fp32_exact": round(rec, 5),
                   "final_abs_err_vs_fp64": {"max": float(err_final.max()), "mean": float(err_final.mean())},
                   "candidate_abs_err_vs_fp64": {"max": float(err_cand.max()), "mean": float(err_cand.mean())},
                   "error_sample": len(sub)},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": peak,
                     "unit": "TFLOP/s", "frac": round(tflops / peak, 4),
                     "traffic": _traffic(f"allknn_{fmt}_1m96_traffic.json", "hbm_bytes_per_call",
                                         n == 1_000_000),
                     "traffic_source": f"profiles/allknn_{fmt}_1m96.md",
                     "kernel": f"prec_stage1<{'3, 8, false' if fmt == 'bf16' else '1, 8, true'}> + "
                               "prec_stage2<6> (whole call)",
                     "kernel_ms": round(ms, 3), "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }


def gather_bench(args):
    """Calibration of the HBM counters on the beam kernels' access shape: rerank! (one
    wave per column, rows gathered by id) over n = 1,000,000 x 128 FP32 rows where the ids
    of all columns together are one random permutation of 1..n (32 per column, 31,250
    columns): every row is read exactly once, at random.  Algorithmic bytes per call:
    n * 512 (rows) + 31,250 * 512 (query rows) + 31,250 * 32 * 8 (ids and distances read)
    + 31,250 * 32 * 8 (written).  Under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) the
    counters' readings per dispatch against these bytes give the factor for random 512 B
    gathers, to set beside the 1 GiB copy's."""
    import numpy as np
    from similaritysearch import _lib, datasets
    from similaritysearch.device import DeviceArray, Event, Stream
    import similaritysearch as ss

    L = _lib.load()
    n, dim, kc = args.n, 128, 32
    m = n // kc
    X = datasets.gmm(n, dim, seed=4)
    Q = datasets.gmm(m, dim, seed=5)
    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    h = E._handle()
    ids = (np.random.default_rng(13).permutation(n)[:m * kc] + 1).astype(np.uint32).reshape(m, kc)
    st = Stream()
    dQ = DeviceArray.from_numpy(Q)
    dI0 = DeviceArray.from_numpy(ids)
    dI = DeviceArray((m, kc), np.uint32)
    dD = DeviceArray((m, kc), np.float32)

    def run():
        _lib.check(L.ssg_memcpy(dI.ptr, dI0.ptr, ids.nbytes, 2), "reset ids")
        _lib.check(L.ssg_rerank_dev(h, dQ.ptr, m, kc, dI.ptr, dD.ptr, st.handle), "rerank!")
    for _ in range(max(1, args.warmup)):
        run()
    st.synchronize()
    e0, e1 = Event(), Event()
    t = 0.0
    for _ in range(args.steps):
        _lib.check(L.ssg_memcpy(dI.ptr, dI0.ptr, ids.nbytes, 2), "reset ids")
        e0.record(st)
        _lib.check(L.ssg_rerank_dev(h, dQ.ptr, m, kc, dI.ptr, dD.ptr, st.handle), "rerank!")
        e1.record(st)
        st.synchronize()
        t += e0.elapsed_ms(e1)
    ms = t / args.steps
    rd = n * dim * 4 + m * dim * 4 + m * kc * 8
    wr = m * kc * 8
    gbs = (rd + wr) / (ms * 1e-3) / 1e9
    log(f"gather: {ms:.3f} ms per rerank of {m} columns x {kc}: {gbs:.0f} GB/s")
    return {"metric": "rerank! random row gather GB/s (each row once)", "value": round(gbs, 1),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f32", "data": "GMM rows, ids = a random permutation",
            "config": {"workload": f"ssg_rerank_dev over {m} columns of {kc} ids, {n}x{dim} FP32",
                       "read_bytes": rd, "write_bytes": wr},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                         "kernel": "rerank_kernel", "kernel_ms": round(ms, 3),
                         "algorithmic_bytes_per_launch": rd + wr},
            "cpu_baseline": None}


def hsp_bench(args):
    """hsp_queries (src/hsp.jl:77-101) of every object of the 1,000,000 x 96 unit-vector
    database (N(0, I) seed 6, L2-normalised), NormCosine, over its k = 32 nearest
    neighbours (allknn with the FP8 candidate pass and FP32 rerank, computed once before
    the timed region).  A step is one ssg_hsp_queries_dev launch over all n columns;
    `value` is columns (queries) per second.  `roofline`: HBM-bound; the algorithmic bytes
    of a column are its k candidate rows (k * dim * 4) plus its id and distance columns
    read and written (16 k), over the launch's HIP-event duration.  Checked bit for bit
    against the oracle on 5,000 evenly spaced columns; `cpu_baseline`: the oracle's
    restated hsp_queries (OpenMP over columns) on a bounded sample of the same columns."""
    import numpy as np
    import similaritysearch as ss
    from similaritysearch import datasets, _lib
    from similaritysearch.device import DeviceArray, Event, Stream

    L = _lib.load()
    n = args.n
    dim, k, ncand = 96, 32, 128
    X = datasets.unit(n, dim, seed=6)
    E = ss.ExhaustiveSearch(ss.NormCosine, ss.MatrixDatabase.from_rows(X))
    h = E._handle()
    st = Stream()
    sp = st.handle
    dI = DeviceArray((n, k), np.uint32)
    dD = DeviceArray((n, k), np.float32)
    dHI = DeviceArray((n, k), np.uint32)
    dHD = DeviceArray((n, k), np.float32)
    _lib.check(L.ssg_allknn_exhaustive_prec_dev(h, k, 2, ncand, dI.ptr, dD.ptr, None, None, sp),
               "allknn")

    def run():
        _lib.check(L.ssg_hsp_queries_dev(h, n, k, dI.ptr, dD.ptr, dHI.ptr, dHD.ptr, sp),
                   "hsp_queries")

    for _ in range(max(1, args.warmup)):
        run()
    st.synchronize()
    e0, e1 = Event(), Event()
    e0.record(st)
    for _ in range(args.steps):
        run()
    e1.record(st)
    st.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    ids, dists = dI.numpy(), dD.numpy()
    hi, hd = dHI.numpy(), dHD.numpy()
    kept = float((hi != 0).sum(1).mean())
    log(f"hsp_queries: {ms:.3f} ms per {n} columns of k={k}; mean kept {kept:.2f}")

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sub = np.linspace(0, n - 1, min(n, 5_000)).astype(np.int64)
    oi, od = pyoracle.hsp_queries("NormCosine", X, None, ids[sub], dists[sub], nthreads=threads)
    ok = bool((oi == hi[sub]).all() and (od == hd[sub]).all())
    log(f"oracle match on {len(sub)} columns: {ok}")

    cpu = None
    if not args.no_cpu:
        nq, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and nq < n:
            pyoracle.hsp_queries("NormCosine", X, None, ids[nq:nq + 20_000], dists[nq:nq + 20_000],
                                 nthreads=threads, baseline=True)
            nq += 20_000
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(min(nq, n) / cpu_s, 1), "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": f"restated hsp_queries (hsp_should_push with the reference's FP32 SIMD "
                         f"distance, KnnSorted) of the first {min(nq, n)} columns, OpenMP over columns"}

    bytes_col = k * dim * 4 + 16 * k
    gbs = n * bytes_col / (ms * 1e-3) / 1e9
    # HBM traffic of this launch from the committed counter measurement of the same kernel
    # and workload (FETCH_SIZE / WRITE_SIZE passes scaled by a 1 GiB copy)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hsp_queries_1m96_k32_traffic.json")
    if os.path.exists(tfile) and n == 1_000_000:
        with open(tfile) as fh:
            traffic = int(json.load(fh)["hbm_bytes_per_column"] * n)
    out = {
        "metric": "hsp_queries columns/s, 1M×96 NormCosine, k=32 candidates per column",
        "value": round(n / (ms * 1e-3), 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: N(0, I) rows, seed 6, L2-normalised; candidates = allknn k=32",
        "config": {"workload": f"hsp_queries over the allknn of {n}x{dim} NormCosine, k={k}",
                   "mean_kept": round(kept, 3), "oracle_match_5000": ok},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/hsp_queries_1m96_k32.md",
                     "kernel": "hsp_kernel<6, 1>", "kernel_ms": round(ms, 3),
                     "bytes_per_launch": n * bytes_col},
        "cpu_baseline": cpu,
    }
    return out


def copy_bench(args):
    """Calibration for the HBM counters: one 1 GiB device-to-device copy (ssg_memcpy) per
    step.  Under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) restricted to the copy
    kernel, the counters' readings for 2^30 bytes moved give the factor that turns a
    kernel's readings into HBM bytes."""
    import numpy as np
    from similaritysearch import _lib
    from similaritysearch.device import DeviceArray, synchronize

    L = _lib.load()
    nb = 1 << 30
    a = DeviceArray((nb // 4,), np.float32)
    b = DeviceArray((nb // 4,), np.float32)
    for _ in range(max(1, args.warmup)):
        _lib.check(L.ssg_memcpy(b.ptr, a.ptr, nb, 2), "copy")
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _lib.check(L.ssg_memcpy(b.ptr, a.ptr, nb, 2), "copy")
    synchronize()
    s_ = (time.perf_counter() - t0) / args.steps
    return ({"metric": "device-to-device copy GB/s (read + write)", "value": round(2 * nb / s_ / 1e9, 1),
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(s_ * 1e3, 3), "higher_is_better": True, "scaling": "strong",
                      "vs_baseline": None, "dtype": "u8", "data": "1 GiB buffer",
                      "config": {"workload": "ssg_memcpy 1 GiB device to device"}})


def neardup_data(n, dim, seed):
    """n unit vectors in clusters of four: n/4 random bases (N(0, I), L2-normalised) and
    three copies of each with N(0, 0.01^2) noise, renormalised, in a seeded random order."""
    import numpy as np
    from similaritysearch import datasets
    nb = n // 4
    B = datasets.unit(nb, dim, seed=seed)
    rng = np.random.default_rng(seed + 1)
    X = np.empty((n, dim), np.float32)
    X[:nb] = B
    for c in range(1, 4):
        Y = B + rng.standard_normal(B.shape, dtype=np.float32) * np.float32(0.01)
        X[c * nb:(c + 1) * nb] = Y / np.linalg.norm(Y, axis=1, keepdims=True)
    return X[rng.permutation(n)]


def neardup_bench(args):
    """neardup(dist, X, eps) with an exact index (src/neardup.jl:69-256) over 1,000,000 unit
    vectors of 96 dims in 250,000 clusters of four near duplicates (NormCosine, eps = 0.05,
    blocksize 256, k = 8).  A step is one whole neardup; `value` is objects per second.
    The net is searched by the exact FP32 exhaustive search on an index that grows by
    append (ssg_append_db), the in-block distances come from rerank!, and the ε-net
    decisions are host-side.  `roofline`: the net searches' 2 * dim * block * |net| MFMA
    flops over the whole call (a floor: host work and the rerank launches are in the
    time).  Checked bit for bit against the oracle on the first 20,000 objects;
    `cpu_baseline`: the oracle's neardup (OpenMP exhaustive search of the net) on a
    prefix."""
    import numpy as np
    import similaritysearch as ss

    n, dim, eps = args.n, 96, 0.05
    X = neardup_data(n, dim, seed=12)
    db = ss.MatrixDatabase.from_rows(X)
    ss.neardup(ss.NormCosine, ss.MatrixDatabase.from_rows(X[:4096]), eps)   # warm-up
    t = []
    R = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        R = ss.neardup(ss.NormCosine, db, eps)
        t.append(time.perf_counter() - t0)
    s_ = float(np.median(t))
    centers = len(R["centers"])
    log(f"neardup: {s_:.2f} s per neardup of {n} ({centers} centers, costdists {R['costdists']})")
    # net-search flops: each block after the first searches the net built so far
    flops, net, bs = 0.0, 0, 256
    cen_pos = np.sort(R["centers"].astype(np.int64))
    for r0 in range(0, n, bs):
        net_before = int(np.searchsorted(cen_pos, r0 + 1))
        if net_before:
            flops += 2.0 * dim * min(bs, n - r0) * net_before
    tflops = flops / s_ / 1e12

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sub = 20_000
    Rs = ss.neardup(ss.NormCosine, ss.MatrixDatabase.from_rows(X[:sub]), eps)
    M, L, D, cen, cost = pyoracle.neardup("NormCosine", X[:sub], eps, nthreads=threads)
    ok = bool((Rs["map"] == M).all() and (Rs["nn"] == L).all() and (Rs["dist"] == D).all()
              and (Rs["centers"] == cen).all() and Rs["costdists"] == cost)
    log(f"oracle match on the first {sub} objects: {ok}")

    cpu = None
    if not args.no_cpu:
        m, t0 = 25_000, time.perf_counter()
        pyoracle.neardup("NormCosine", X[:m], eps, nthreads=threads)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(m / cpu_s, 1), "unit": "objects/s", "cores": threads, "kind": "port",
               "sample": f"the oracle's neardup of the first {m} objects (exhaustive net searches "
                         f"with OpenMP, block filter on the host); the net grows with n, so a "
                         f"whole neardup runs slower per object than this prefix"}
    out = {
        "metric": "neardup objects/s, 1M×96 NormCosine, eps=0.05, blocksize 256",
        "value": round(n / s_, 1), "unit": "objects/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(s_ * 1e3, 1), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: 250,000 N(0, I) unit bases x 4 noisy copies (sigma 0.01), shuffled, seed 12",
        "config": {"workload": f"neardup with an exact index over {n}x{dim} NormCosine, eps={eps}",
                   "centers": centers, "costdists": int(R["costdists"]), "oracle_match_20000": ok},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(tflops / 157.3, 5), "traffic": None,
                     "kernel": "exh_stage1 net searches (whole call, host work included)",
                     "kernel_ms": round(s_ * 1e3, 1), "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }
    return out


if __name__ == "__main__":
    main()
