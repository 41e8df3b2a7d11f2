"""Inference with the fused output layer (head_top_kernel: the logits are folded into the
landmark top-h, softmax sum and bucket logits without being stored) against the stored-logits
path (output-layer GEMM, then topk_heads_kernel), which ssg_mlp_infer_dev takes for h > 4.
Both form the logits with the same products, so the landmarks and buckets are equal bit for
bit; the confidence (1 / softmax sum of the landmark head) is summed in another order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _infer(L, mlp, dQ, m, land, h, tau):
    from similaritysearch import _lib
    from similaritysearch.device import DeviceArray, Stream
    st = Stream()
    dH, dB, dC = DeviceArray((m, h), np.uint32), DeviceArray((m,), np.int32), DeviceArray((m,), np.float32)
    _lib.check(L.ssg_mlp_infer_dev(mlp._p, dQ.ptr, m, land.ptr, h, float(tau), dH.ptr, dB.ptr, dC.ptr,
                                   st.handle), "infer")
    st.synchronize()
    return dH.numpy(), dB.numpy(), dC.numpy()


@pytest.mark.parametrize("hidden", [64, 128, 192, 256])
@pytest.mark.parametrize("tau", [0.0, 0.65])
def test_fused_inference_equals_stored_logits(hidden, tau):
    from similaritysearch import _lib
    from similaritysearch.learned import Mlp
    from similaritysearch.device import DeviceArray
    L = _lib.load()
    D, H1, NL, NB, m = 48, 80, 301, 8, 777          # 301 landmarks: a partial last unit tile
    mlp = Mlp(D, H1, hidden, NL, NB, seed=hidden)
    rng = np.random.default_rng(hidden)
    Q = rng.standard_normal((m, D)).astype(np.float32)
    dQ = DeviceArray.from_numpy(Q)
    # landmark ids in shuffled order, so a wrong unit index shows as a wrong id
    land = DeviceArray.from_numpy((rng.permutation(NL) + 1).astype(np.uint32))
    H5, B5, C5 = _infer(L, mlp, dQ, m, land, 5, tau)
    print(f"hidden {hidden} tau {tau}: confidence range {C5.min():.4f} .. {C5.max():.4f}, "
          f"buckets {np.bincount(B5, minlength=NB).tolist()}")
    for h in (1, 2, 3, 4):
        Hh, Bh, Ch = _infer(L, mlp, dQ, m, land, h, tau)
        np.testing.assert_array_equal(Hh, H5[:, :h])
        np.testing.assert_array_equal(Bh, B5)
        np.testing.assert_allclose(Ch, C5, rtol=2e-6, atol=0)


def test_fused_inference_large_batch_and_bucket_head_tile():
    """More rows than one 128-query block per CU round, and 12 buckets whose logits straddle
    two unit tiles (4090 landmarks: units 4090..4101)."""
    from similaritysearch import _lib
    from similaritysearch.learned import Mlp
    from similaritysearch.device import DeviceArray
    L = _lib.load()
    D, NL, NB, m = 128, 4090, 12, 70_001
    mlp = Mlp(D, 256, 256, NL, NB, seed=5)
    rng = np.random.default_rng(5)
    dQ = DeviceArray.from_numpy(rng.standard_normal((m, D)).astype(np.float32))
    land = DeviceArray.from_numpy(np.arange(1, NL + 1, dtype=np.uint32))
    H5, B5, C5 = _infer(L, mlp, dQ, m, land, 5, 0.5)
    H2, B2, C2 = _infer(L, mlp, dQ, m, land, 2, 0.5)
    np.testing.assert_array_equal(H2, H5[:, :2])
    np.testing.assert_array_equal(B2, B5)
    np.testing.assert_allclose(C2, C5, rtol=2e-6, atol=0)
