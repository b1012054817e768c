"""The MLP oracle's backward is pinned by central finite differences."""
import numpy as np

from oracle import mlp


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init_params(rng)
    X = rng.normal(size=(7, mlp.NIN))
    T = rng.normal(size=(7, mlp.NOUT))

    def loss(q):
        Y, _, _ = mlp.forward(q, X, bf16=False)
        return 0.5 * np.sum((Y - T) ** 2)

    Y, H1, H2 = mlp.forward(p, X, bf16=False)
    g = mlp.backward(p, X, H1, H2, Y - T, bf16=False)
    eps = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            q = {kk: vv.copy() for kk, vv in p.items()}
            q[k].reshape(-1)[idx] += eps
            lp = loss(q)
            q[k].reshape(-1)[idx] -= 2 * eps
            lm = loss(q)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[k].reshape(-1)[idx]) <= 1e-5 * max(1.0, abs(fd)), (k, idx)


def test_bf16_round_is_rne():
    x = np.array([1.0, 1.00390625, 1.0078125, 1.005859375, -3.14159, 65504.0], np.float32)
    r = mlp.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125    # tie -> even
    assert r[3] == 1.0078125
    assert abs(r[4] + 3.140625) < 1e-9
