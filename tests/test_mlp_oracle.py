"""The numpy restatement of the predictor network (oracle/mlp_oracle.py)
pinned by finite differences and by a hand-derivable case."""
import numpy as np

import mlp_oracle as mo


def random_net(dims, seed):
    rng = np.random.default_rng(seed)
    Ws = [rng.uniform(-1, 1, (dims[l + 1], dims[l])) / np.sqrt(dims[l]) for l in range(len(dims) - 1)]
    bs = [rng.uniform(-0.1, 0.1, dims[l + 1]) for l in range(len(dims) - 1)]
    return Ws, bs


def test_backward_matches_finite_differences():
    dims = [7, 5, 6, 3]
    Ws, bs = random_net(dims, 0)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((11, 7))
    T = rng.standard_normal((11, 3))

    def loss_of(Ws, bs):
        Y, _ = mo.forward(Ws, bs, X, round_bf16=False)
        return mo.loss_and_grad(Y, T, round_bf16=False)[0]

    Y, acts = mo.forward(Ws, bs, X, round_bf16=False)
    _, dY = mo.loss_and_grad(Y, T, round_bf16=False)
    gW, gb = mo.backward(Ws, acts, dY, round_bf16=False)
    eps = 1e-6
    for l in range(len(Ws)):
        for (i, j) in [(0, 0), (1, 2), (dims[l + 1] - 1, dims[l] - 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][i, j] += eps
            Wm[l][i, j] -= eps
            fd = (loss_of(Wp, bs) - loss_of(Wm, bs)) / (2 * eps)
            assert abs(fd - gW[l][i, j]) <= 1e-8 + 1e-6 * abs(fd)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][0] += eps
        bm[l][0] -= eps
        fd = (loss_of(Ws, bp) - loss_of(Ws, bm)) / (2 * eps)
        assert abs(fd - gb[l][0]) <= 1e-8 + 1e-6 * abs(fd)


def test_single_linear_layer_closed_form():
    # one linear layer: Y = X W^T + b, L = mean((Y-T)^2) -> dW = 2/n (Y-T)^T X
    rng = np.random.default_rng(2)
    W = rng.standard_normal((2, 3))
    b = rng.standard_normal(2)
    X = rng.standard_normal((4, 3))
    T = rng.standard_normal((4, 2))
    Y, acts = mo.forward([W], [b], X, round_bf16=False)
    np.testing.assert_allclose(Y, X @ W.T + b)
    _, dY = mo.loss_and_grad(Y, T, round_bf16=False)
    gW, gb = mo.backward([W], acts, dY, round_bf16=False)
    np.testing.assert_allclose(gW[0], 2.0 / Y.size * (Y - T).T @ X, rtol=1e-14)
    np.testing.assert_allclose(gb[0], 2.0 / Y.size * (Y - T).sum(axis=0), rtol=1e-14)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], dtype=np.float64)
    r = mo.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                 # tie -> even
    assert r[2] == 1.0 + 2 ** -6       # tie -> even (upwards)
    assert r[3] == -2.5
    assert abs(r[4] - 1e-3) <= 2 ** -9 * 1e-3
