"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
oracle): the oracle must keep reproducing them (CPU), and the HIP path must reproduce
them bit-exactly through the C ABI (GPU)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ----------------------------------------------------------------------- oracle (CPU)
@pytest.mark.parametrize("name,metric", [("exhaustive_sql2_2000x32.npz", "SqL2"),
                                         ("exhaustive_normcos_1500x24.npz", "NormCosine")])
def test_oracle_exhaustive_fixture(oracle, name, metric):
    z = load(name)
    ids, d = oracle.exhaustive(metric, z["X"], z["Q"], int(z["k"]))
    np.testing.assert_array_equal(ids, z["ids"])
    np.testing.assert_array_equal(d, z["dists"])


def test_oracle_beam_fixture(oracle):
    z = load("beam_gmm_2500x16.npz")
    for tag in ("a", "b"):
        k, bs, delta, mv = z[f"params_{tag}"]
        ids, d, cd, cb = oracle.beamsearch("SqL2", z["X"], z["offs"], z["adj"], z["hints"], z["Q"],
                                           int(k), int(bs), float(delta), int(mv))
        np.testing.assert_array_equal(ids, z[f"ids_{tag}"])
        np.testing.assert_array_equal(d, z[f"dists_{tag}"])
        np.testing.assert_array_equal(cd, z[f"costd_{tag}"])
        np.testing.assert_array_equal(cb, z[f"costb_{tag}"])


def test_oracle_builder_fixture(oracle):
    z = load("beam_gmm_2500x16.npz")
    g = oracle.build_graph("SqL2", z["X"], parallel_block=1, seed=5)
    np.testing.assert_array_equal(g["offs"], z["offs"])
    np.testing.assert_array_equal(g["adj"], z["adj"])
    np.testing.assert_array_equal(g["hints"], z["hints"])


def test_oracle_knnsorted_fixture(oracle):
    z = load("knnsorted_script.npz")
    ids, d, acc = oracle.knn_run(int(z["cap"]), z["ops"], z["ids_in"], z["d_in"])
    np.testing.assert_array_equal(ids, z["ids"])
    np.testing.assert_array_equal(d, z["dists"])
    np.testing.assert_array_equal(acc, z["accepted"].astype(bool))


def test_oracle_fft_fixture(oracle):
    z = load("fft_600x8.npz")
    c, nn, nnd, eps = oracle.fft("SqL2", z["X"], int(z["k"]), start=int(z["start"]))
    for a, b in ((c, "centers"), (nn, "nn"), (nnd, "nnd"), (eps, "eps")):
        np.testing.assert_array_equal(a, z[b])


def test_oracle_rerank_fixture(oracle):
    z = load("rerank_400x20.npz")
    ids, d = oracle.rerank("SqL2", z["X"], z["Q"], z["cand"], np.full(z["cand"].shape, np.inf, np.float32))
    np.testing.assert_array_equal(ids, z["ids"])
    np.testing.assert_array_equal(d, z["dists"])


# ----------------------------------------------------------------------- HIP path (GPU)
@pytest.mark.gpu
@pytest.mark.parametrize("name,metric", [("exhaustive_sql2_2000x32.npz", "SqL2"),
                                         ("exhaustive_normcos_1500x24.npz", "NormCosine")])
def test_gpu_exhaustive_fixture(name, metric):
    import similaritysearch as ss
    z = load(name)
    E = ss.ExhaustiveSearch(getattr(ss, metric), ss.MatrixDatabase.from_rows(z["X"]))
    ids, d = ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(z["Q"]), int(z["k"]))
    np.testing.assert_array_equal(ids.T, z["ids"])
    np.testing.assert_array_equal(d.T, z["dists"])


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [1, 0])
def test_gpu_beam_fixture(layout):
    import ctypes as C
    import similaritysearch as ss
    from similaritysearch import _lib
    z = load("beam_gmm_2500x16.npz")
    G = ss.SearchGraph(ss.SqL2, ss.MatrixDatabase.from_rows(z["X"]),
                       ss.StaticAdjList.from_csr(z["offs"], z["adj"]), hints=z["hints"])
    _lib.check(_lib.load().ssg_set_option(G._handle(), _lib.SSG_OPT_QUERY_LAYOUT, layout), "option")
    ctx = ss.SearchGraphContext()
    for tag in ("a", "b"):
        k, bs, delta, mv = z[f"params_{tag}"]
        G.algo = ss.BeamSearch(int(bs), float(delta), int(mv))
        ids, d = ss.searchbatch(G, ctx, ss.MatrixDatabase.from_rows(z["Q"]), int(k))
        np.testing.assert_array_equal(ids.T, z[f"ids_{tag}"])
        np.testing.assert_array_equal(d.T, z[f"dists_{tag}"])
        np.testing.assert_array_equal(ctx.costdists, z[f"costd_{tag}"])
        np.testing.assert_array_equal(ctx.costblocks, z[f"costb_{tag}"])


@pytest.mark.gpu
def test_gpu_fft_fixture():
    import similaritysearch as ss
    z = load("fft_600x8.npz")
    c, nn, nnd, eps = ss.fft(ss.SqL2, ss.MatrixDatabase.from_rows(z["X"]), int(z["k"]), start=int(z["start"]))
    for a, b in ((c, "centers"), (nn, "nn"), (nnd, "nnd"), (eps, "eps")):
        np.testing.assert_array_equal(a, z[b])


@pytest.mark.gpu
def test_gpu_rerank_fixture():
    import similaritysearch as ss
    z = load("rerank_400x20.npz")
    ids = np.asfortranarray(z["cand"].T.copy())
    d = np.asfortranarray(np.full(ids.shape, np.inf, np.float32))
    ss.rerank_(ss.SqL2, ss.MatrixDatabase.from_rows(z["X"]), ss.MatrixDatabase.from_rows(z["Q"]), ids, d)
    np.testing.assert_array_equal(ids.T, z["ids"])
    np.testing.assert_array_equal(d.T, z["dists"])
