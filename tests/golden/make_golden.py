"""Writes the golden fixtures of this directory from the CPU oracle (oracle/).

The reference ships no fixtures for this path and its Julia toolchain is not available,
so the vectors come from the oracle, whose own pins are listed in the test suite
(tests/test_oracle.py, tests/test_oracle_mlp.py).  Each .npz holds the inputs and the
expected outputs as plain arrays (loadable with numpy.load(..., allow_pickle=False)).
Regenerate with:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle  # noqa: E402


def gmm(n, dim, centers, sigma, rng):
    C = rng.uniform(-1, 1, (centers, dim)).astype(np.float32)
    lab = rng.integers(0, centers, n)
    return (C[lab] + sigma * rng.standard_normal((n, dim))).astype(np.float32)


def main():
    # exhaustive, SqL2: the reference benchmark's uniform [0, 1) shape at reduced size
    rng = np.random.default_rng(1)
    X = rng.random((2000, 32), dtype=np.float32)
    Q = np.random.default_rng(2).random((100, 32), dtype=np.float32)
    ids, d = pyoracle.exhaustive("SqL2", X, Q, 10)
    np.savez(os.path.join(HERE, "exhaustive_sql2_2000x32.npz"), X=X, Q=Q, k=10, ids=ids, dists=d)

    # exhaustive, NormCosine on unit vectors, odd dim and k
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1500, 24)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    Q = rng.standard_normal((50, 24)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    ids, d = pyoracle.exhaustive("NormCosine", X, Q, 7)
    np.savez(os.path.join(HERE, "exhaustive_normcos_1500x24.npz"), X=X, Q=Q, k=7, ids=ids, dists=d)

    # beam search on a graph built by the restated index! (sequential insertion)
    rng = np.random.default_rng(4)
    X = gmm(2500, 16, 12, 0.2, rng)
    Q = gmm(60, 16, 12, 0.2, np.random.default_rng(5))
    g = pyoracle.build_graph("SqL2", X, parallel_block=1, seed=5)
    out = dict(X=X, Q=Q, offs=g["offs"], adj=g["adj"], hints=g["hints"],
               build_bsize=g["bsize"], build_delta=g["delta"], build_maxvisits=g["maxvisits"])
    for tag, (k, bs, delta, mv) in {"a": (10, 8, 1.0, 10**6), "b": (5, 4, 1.1, 60)}.items():
        ids, d, cd, cb = pyoracle.beamsearch("SqL2", X, g["offs"], g["adj"], g["hints"], Q, k, bs,
                                             delta, mv)
        out.update({f"params_{tag}": np.array([k, bs, delta, mv], np.float64),
                    f"ids_{tag}": ids, f"dists_{tag}": d, f"costd_{tag}": cd, f"costb_{tag}": cb})
    np.savez(os.path.join(HERE, "beam_gmm_2500x16.npz"), **out)

    # KnnSorted: a fixed push/pop script with ties and a full queue
    rng = np.random.default_rng(6)
    ops = (rng.random(400) < 0.15).astype(np.int8)
    ops[:12] = 0
    ids_in = rng.integers(1, 50, 400).astype(np.uint32)
    d_in = np.round(rng.random(400) * 8).astype(np.float32)  # many equal distances
    oi, od, acc = pyoracle.knn_run(10, ops, ids_in, d_in)
    np.savez(os.path.join(HERE, "knnsorted_script.npz"), cap=10, ops=ops, ids_in=ids_in, d_in=d_in,
             ids=oi, dists=od, accepted=acc.astype(np.uint8))

    # fft k-centers from object 1
    rng = np.random.default_rng(7)
    X = gmm(600, 8, 5, 0.3, rng)
    centers, nn, nnd, eps = pyoracle.fft("SqL2", X, 20, start=1)
    np.savez(os.path.join(HERE, "fft_600x8.npz"), X=X, k=20, start=1, centers=centers, nn=nn,
             nnd=nnd, eps=eps)

    # rerank! of shuffled candidate columns with a zero-id tail
    rng = np.random.default_rng(8)
    X = rng.random((400, 20), dtype=np.float32)
    Q = rng.random((30, 20), dtype=np.float32)
    cand = np.zeros((30, 12), np.uint32)
    for j in range(30):
        ln = rng.integers(5, 13)
        cand[j, :ln] = rng.choice(400, ln, replace=False) + 1
    d0 = np.full((30, 12), np.inf, np.float32)
    ids, d = pyoracle.rerank("SqL2", X, Q, cand, d0)
    np.savez(os.path.join(HERE, "rerank_400x20.npz"), X=X, Q=Q, cand=cand, ids=ids, dists=d)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
