"""HIP MLP forward/backward (BF16 MFMA, FP32 accumulate) against the NumPy
oracle with the same BF16 operand rounding; also pins the MFMA lane layout."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [32, 77, 1000])
def test_forward_backward_match_oracle(n):
    import torch
    from rlmusclev2_amd.mlp import Mlp
    rng = np.random.default_rng(n)
    p = omlp.init_params(rng)
    net = Mlp(p)
    pq = net.numpy_params()          # float32-rounded master weights
    X = rng.normal(size=(n, omlp.NIN)).astype(np.float32)
    X[:, 22:] = 0.0
    Xt = torch.tensor(X, device="cuda")
    Y, H1, H2 = net.forward(Xt, keep=True)
    torch.cuda.synchronize()
    # layer by layer, each fed the GPU's previous activation so BF16 rounding sees the same inputs
    h1, h2 = H1.cpu().numpy().astype(np.float64), H2.cpu().numpy().astype(np.float64)
    H1r = np.maximum(omlp.bf16_round(X) @ omlp.bf16_round(pq["W1"]).T + pq["b1"], 0)
    H2r = np.maximum(omlp.bf16_round(h1) @ omlp.bf16_round(pq["W2"]).T + pq["b2"], 0)
    Yr = h2 @ pq["W3"].T + pq["b3"]
    np.testing.assert_allclose(h1, H1r, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(h2, H2r, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(Y.cpu().numpy(), Yr, rtol=1e-5, atol=1e-5)
    # end to end against the oracle forward (BF16 rounding of recomputed activations may differ by 1 ulp)
    Yo, _, _ = omlp.forward(pq, X)
    np.testing.assert_allclose(Y.cpu().numpy(), Yo, rtol=2e-2, atol=2e-2)
    dY = rng.normal(size=(n, omlp.NOUT)).astype(np.float32)
    g = net.backward(Xt, H1, H2, torch.tensor(dY, device="cuda"))
    gr = omlp.backward(pq, X, H1.cpu().numpy().astype(np.float64), H2.cpu().numpy().astype(np.float64),
                       dY.astype(np.float64))
    for k in gr:
        got = g[k].cpu().numpy()
        scale = max(1e-6, np.abs(gr[k]).max())
        err = np.abs(got - gr[k]).max() / scale
        assert err < 2e-3, (k, err)


def test_training_reduces_loss_on_gpu():
    import torch
    from rlmusclev2_amd.mlp import Mlp
    rng = np.random.default_rng(0)
    X = rng.normal(size=(4096, omlp.NIN)).astype(np.float32)
    X[:, 22:] = 0
    T = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), X[:, 2] * X[:, 3], 0.3 * X[:, 4], 0 * X[:, 0]], 1)
    Xt, Tt = torch.tensor(X, device="cuda"), torch.tensor(T.astype(np.float32), device="cuda")
    net = Mlp(seed=1)
    losses = []
    for it in range(300):
        Y, H1, H2 = net.forward(Xt, keep=True)
        dY = (Y - Tt) / X.shape[0]
        losses.append(float(0.5 * ((Y - Tt) ** 2).sum() / X.shape[0]))
        net.adam_step(net.backward(Xt, H1, H2, dY), lr=3e-3)
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
