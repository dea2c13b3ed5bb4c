"""The CPU FP64 rank model (oracle/mlp.py) pinned before it checks the GPU
kernels: gradients against central finite differences, the ReLU mask and
the Adam step against their closed forms."""
import numpy as np

from oracle import mlp


def data(seed, B=96):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, mlp.N_IN))
    y = rng.uniform(0, 6, size=B)
    return X, y


def test_layout():
    assert mlp.N_PARAMS == 64 * 16 + 64 + 2 * (64 * 64 + 64) + 64 + 1 == 9473
    p = np.arange(mlp.N_PARAMS, dtype=np.float64)
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = mlp.unpack(p)
    assert W1[0, 1] == 1 and b1[0] == 1024 and W2[0, 0] == 1088 and b4[0] == 9472
    assert W4.shape == (1, 64)


def test_gradient_matches_finite_differences():
    X, y = data(1)
    p = mlp.init_params(2)
    p[1024:1088] = 0.05  # nonzero biases: every term of the gradient takes part
    _, g = mlp.loss_and_grad(p, X, y)
    rng = np.random.default_rng(3)
    # a sample of every parameter block
    blocks = [(0, 1024), (1024, 1088), (1088, 5184), (5184, 5248), (5248, 9344), (9344, 9408),
              (9408, 9472), (9472, 9473)]
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(12, b - a), replace=False)
                          for a, b in blocks])
    fd = mlp.finite_difference_grad(p, X, y, idx)
    err = np.abs(fd - g[idx]) / np.maximum(1e-6, np.abs(g[idx]))
    assert err.max() <= 1e-5, (err.max(), idx[np.argmax(err)])


def test_a_planted_transpose_fails_the_check():
    """The check above catches a plausible mistake: dX computed with W
    transposed the wrong way round (here on a square layer)."""
    X, y = data(4)
    p = mlp.init_params(5)
    layers = mlp.unpack(p)
    out, acts = mlp.forward(p, X)
    d = (2.0 / len(out)) * (out - y)
    dz = d[:, None]
    W4 = layers[3][0]
    dz = (dz @ W4) * (acts[3] > 0)
    W3 = layers[2][0]
    good = (dz @ W3) * (acts[2] > 0)
    bad = (dz @ W3.T) * (acts[2] > 0)
    gW2_good = good.T @ acts[1]
    gW2_bad = bad.T @ acts[1]
    _, g = mlp.loss_and_grad(p, X, y)
    np.testing.assert_allclose(g[1088:5184].reshape(64, 64), gW2_good, rtol=1e-10, atol=1e-14)
    assert np.abs(gW2_bad - gW2_good).max() > 1e-3 * np.abs(gW2_good).max()


def test_adam_step_closed_form():
    rng = np.random.default_rng(6)
    p = rng.standard_normal(50)
    g = rng.standard_normal(50)
    m = np.zeros(50)
    v = np.zeros(50)
    p1, m1, v1 = mlp.adam_step(p, g, m, v, 1, 1e-3)
    # first step: m_hat = g, v_hat = g^2, so the update is lr * sign(g) (eps aside)
    np.testing.assert_allclose(p - p1, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(v1, 0.001 * g * g)


def test_training_lowers_the_loss():
    rng = np.random.default_rng(7)
    X = rng.standard_normal((512, mlp.N_IN))
    y = 2.0 + X[:, 0] - 0.5 * X[:, 3] ** 2
    p = mlp.init_params(8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0 = mlp.loss(p, X, y)
    for t in range(1, 201):
        _, g = mlp.loss_and_grad(p, X, y)
        p, m, v = mlp.adam_step(p, g, m, v, t, 3e-3)
    assert mlp.loss(p, X, y) < 0.1 * l0
