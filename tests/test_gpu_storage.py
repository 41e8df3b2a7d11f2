"""GPU indexes over on-disk data: an MMapMatrixDatabase uploads straight from its
mapping, and a SearchGraph saved with save_index and reloaded over a reopened file
searches exactly like the original."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ss():
    import similaritysearch as ss
    return ss


def test_exhaustive_over_mmap_database(tmp_path, oracle):
    ss = _ss()
    rng = np.random.default_rng(31)
    X = rng.random((5000, 40), dtype=np.float32)
    Q = rng.random((300, 40), dtype=np.float32)
    db = ss.MMapMatrixDatabase(str(tmp_path / "x.mmapdb"), 40, capacity_bits=6)
    db.append_items(X[:3000])
    for r in X[3000:3010]:
        db.push_item(r)
    db.append_items(X[3010:])
    E = ss.ExhaustiveSearch(ss.SqL2, db)
    ids, dists = ss.searchbatch(E, ss.GenericContext(), ss.MatrixDatabase.from_rows(Q), 10)
    gi, gd = oracle.exhaustive("SqL2", X, Q, 10)
    np.testing.assert_array_equal(ids.T, gi)
    np.testing.assert_array_equal(dists.T, gd)
    E.close()
    db.close()


def test_saved_graph_searches_like_the_original(tmp_path):
    ss = _ss()
    rng = np.random.default_rng(32)
    X = rng.random((20000, 16), dtype=np.float32)
    Q = ss.MatrixDatabase.from_rows(rng.random((500, 16), dtype=np.float32))
    db = ss.MMapMatrixDatabase(str(tmp_path / "g.mmapdb"), 16)
    db.append_items(X)
    db.close()
    db = ss.MMapMatrixDatabase(str(tmp_path / "g.mmapdb"), read_only=True)
    G = ss.SearchGraph.build(ss.SqL2, db, seed=4)
    ctx = ss.SearchGraphContext()
    ids, dists = ss.searchbatch(G, ctx, Q, 10)
    cd = ctx.costdists.copy()
    ss.save_index(str(tmp_path / "g.ssgidx"), G)
    G.close()
    H = ss.load_index(str(tmp_path / "g.ssgidx"), ss.SqL2,
                      ss.MMapMatrixDatabase(str(tmp_path / "g.mmapdb"), read_only=True))
    ctx2 = ss.SearchGraphContext()
    ids2, dists2 = ss.searchbatch(H, ctx2, Q, 10)
    np.testing.assert_array_equal(ids2, ids)
    np.testing.assert_array_equal(dists2, dists)
    np.testing.assert_array_equal(ctx2.costdists, cd)
    H.close()
