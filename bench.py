#!/usr/bin/env python3
"""Benchmark: QPS at recall@10 >= 0.95 of SearchGraph beam search on one MI355X.

Workload (synthetic, seeded): 1,000,000 x 128 Float32 database drawn from a
100-component isotropic Gaussian mixture (centers ~ U[-1,1]^128, sigma 0.15, equal
weights) and 10,000 queries from the same mixture; squared L2; k = 10.  The graph
is built on the GPU by index! with the reference's default context knobs
(parallel_block = 8 x host threads = 128), autotuned with optimize_index!
(MinRecall(0.9)), then tuned again until recall@10 against the exhaustive ground
truth of the 10,000 queries is >= 0.95.  A step is one searchbatch! over the
10,000 device-resident queries.

Prints one JSON line.  `roofline` is HBM-bound: algorithmic bytes of the timed
launches over their HIP-event duration.  `cpu_baseline` is the oracle's restated
beam search (OpenMP over queries, all host threads) on the same graph and
parameters over the full query set.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=10_000)
    ap.add_argument("--parallel-block", type=int, default=128)
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--params", type=str, default="",
                    help="bsize,delta,maxvisits: skip the recall calibration")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the query-sharded "
              "multi-GPU layout is described in DESIGN.md and not run here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import similaritysearch as ss
    from similaritysearch import datasets, _lib

    dim, k = 128, 10
    t0 = time.time()
    X = datasets.gmm(args.n, dim, seed=4)
    Q = datasets.gmm(args.queries, dim, seed=5)
    Qtune = datasets.gmm(512, dim, seed=6)
    log(f"data {args.n}x{dim} + {args.queries} queries in {time.time() - t0:.1f}s")

    t0 = time.time()
    ctx = ss.SearchGraphContext()
    G = ss.SearchGraph.build(ss.SqL2, ss.MatrixDatabase.from_rows(X), ctx,
                             parallel_block=args.parallel_block, seed=1, verbose=10)
    build_s = time.time() - t0
    log(f"index! {build_s:.1f}s nnz={len(G.adj.end_point)} hints={len(G.hints)} {G.algo}")

    E = ss.ExhaustiveSearch(ss.SqL2, ss.MatrixDatabase.from_rows(X))
    gold, _ = ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(Q), k)
    E.close()

    Qdb = ss.MatrixDatabase.from_rows(Q)

    def run(algo):
        G.algo = algo
        ids, _ = ss.searchbatch(G, ctx, Qdb, k)
        return ss.macrorecall(gold, ids)

    if args.params:
        b, d, mv = args.params.split(",")
        algo = ss.BeamSearch(int(b), float(d), int(mv))
        recall = run(algo)
    else:
        # optimize_index!(MinRecall(0.9)) as the reference tunes the graph, then raise the
        # target until the held-out recall reaches the benchmark's level
        algo, _ = G.optimize(ctx, 0.9, ss.MatrixDatabase.from_rows(Qtune))
        recall = run(algo)
        log(f"MinRecall(0.9): {algo} recall@10={recall:.4f}")
        target = args.target_recall
        while recall < args.target_recall and target < 1.0:
            algo, _ = G.optimize(ctx, min(target, 0.999), ss.MatrixDatabase.from_rows(Qtune))
            recall = run(algo)
            log(f"MinRecall({target:.3f}): {algo} recall@10={recall:.4f}")
            target += 0.01
        if recall < args.target_recall:
            bs = algo.bsize
            while recall < args.target_recall and bs < 256:
                bs *= 2
                algo = ss.BeamSearch(bs, algo.delta, 10**6)
                recall = run(algo)
                log(f"widened: {algo} recall@10={recall:.4f}")
    G.algo = algo
    costd = ctx.costdists.copy()
    costb = ctx.costblocks.copy()

    # device-resident timed loop
    dev = torch.device("cuda:0")
    m = args.queries
    dQ = torch.from_numpy(Q).to(dev)
    dI = torch.zeros((m, k), dtype=torch.int32, device=dev)
    dD = torch.zeros((m, k), dtype=torch.float32, device=dev)
    # a dedicated stream (default priority): the kernels and the timing events share it
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    L = _lib.load()
    h = G._handle()
    bp = G.beam_params(ctx)

    def step():
        _lib.check(L.ssg_beamsearch_dev(h, dQ.data_ptr(), m, k, k, C.byref(bp), None,
                                        dI.data_ptr(), dD.data_ptr(), None, None,
                                        C.c_void_p(stream.cuda_stream)), "searchbatch!")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))
    ids_dev = dI.cpu().numpy().view(np.uint32)
    rec_dev = ss.macrorecall(gold, ids_dev.T)

    qps = m * args.steps / wall
    n = args.n
    deg = len(G.adj.end_point) / n
    # algorithmic bytes per query: E vectors (dim*4) + E visited bytes + B node entries
    # (8 B) + B adjacency lists (4 B x mean degree) + the query (dim*4) + the result (8k)
    bytes_q = (costd * (dim * 4 + 1) + costb * (8 + 4 * deg)).astype(np.float64) + dim * 4 + 8 * k
    step_bytes = float(bytes_q.sum())
    achieved = step_bytes / (kern_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        threads = os.cpu_count() or 1
        offs = np.concatenate([[0], G.adj.offset]).astype(np.int64)
        pyoracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q[:200], k, algo.bsize,
                            algo.delta, algo.maxvisits, nthreads=threads)
        t0 = time.perf_counter()
        reps = 0
        while True:
            ci, _, _, _ = pyoracle.beamsearch("SqL2", X, offs, G.adj.end_point, G.hints, Q, k,
                                              algo.bsize, algo.delta, algo.maxvisits,
                                              nthreads=threads)
            reps += 1
            if time.perf_counter() - t0 > 10.0 or reps >= 20:
                break
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(m * reps / cpu_s, 1), "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": f"restated searchbatch over all {m} queries x{reps}, same graph and "
                         f"BeamSearch, OpenMP over queries; recall@10 "
                         f"{ss.macrorecall(gold, ci.T):.4f}"}

    out = {
        "metric": "QPS at recall@10=0.95, 1M×128 L2, 1 GPU: learned+BF16 vs plain beam search",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 100-component Gaussian mixture, sigma 0.15",
        "config": {"workload": "plain SearchGraph beam search, 1,000,000x128 GMM, 10,000 queries, "
                               "SqL2, k=10",
                   "form": "plain FP32 beam search",
                   "n": n, "queries": m, "k": k, "recall@10": round(rec_dev, 4),
                   "bsize": algo.bsize, "delta": algo.delta, "maxvisits": algo.maxvisits,
                   "beamcap": ctx.beamcap, "parallel_block": args.parallel_block,
                   "graph_nnz": int(len(G.adj.end_point)), "hints": int(len(G.hints)),
                   "build_s": round(build_s, 1),
                   "mean_costdists": round(float(costd.mean()), 1),
                   "mean_costblocks": round(float(costb.mean()), 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "beam_kernel<8,1>", "kernel_ms": round(kern_ms, 3),
                     "algorithmic_bytes_per_launch": int(step_bytes)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
