"""Seeded synthetic workloads (numpy), returned as (n, dim) float32 row arrays
(wrap them with MatrixDatabase.from_rows).

- uniform:  i.i.d. U[0,1) (the exhaustive-search workload, 100,000 x 32).
- gmm:      100-component isotropic Gaussian mixture, centers ~ U[-1,1]^dim,
            sigma 0.15, equal weights (the graph workload, 1,000,000 x 128).
- unit:     standard Gaussian rows normalised to unit L2 norm (the cosine
            workload, 1,000,000 x 96).
"""
import numpy as np


def uniform(n, dim, seed):
    return np.random.default_rng(seed).random((n, dim), dtype=np.float32)


def gmm_centers(dim, ncenters=100, seed=3):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, (ncenters, dim)).astype(np.float32)


def gmm(n, dim, seed, ncenters=100, sigma=0.15, center_seed=3, chunk=1 << 18):
    C = gmm_centers(dim, ncenters, center_seed)
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = rng.integers(0, ncenters, e - s)
        out[s:e] = C[lab] + sigma * rng.standard_normal((e - s, dim), dtype=np.float32)
    return out


def unit(n, dim, seed, chunk=1 << 18):
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = rng.standard_normal((e - s, dim), dtype=np.float32)
        out[s:e] = x / np.linalg.norm(x, axis=1, keepdims=True)
    return out
