"""CPU FP64 restatement of the rank model: 16 inputs, three hidden layers of
64 ReLU units, one linear output, mean squared error, Adam.

TEST INFRASTRUCTURE ONLY: the checker of the library's hand-written training
and inference kernels (forward values, gradients, optimizer steps).  The
reference has no learned component, so nothing here follows a reference
file; the gradients are pinned by central finite differences
(tests/test_oracle_mlp.py) and the optimizer step by its closed form.

Parameter vector layout (float32 on the device, float64 here), row-major
weights W[out][in]:  W1 (64x16), b1 (64), W2 (64x64), b2, W3 (64x64), b3,
W4 (1x64), b4 (1): 9473 values.
"""
import numpy as np

N_IN, N_HID = 16, 64
LAYERS = ((N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, 1))
N_PARAMS = sum(o * i + o for i, o in LAYERS)


def unpack(p):
    """[(W, b)] views of the flat parameter vector, W of shape (out, in)."""
    p = np.asarray(p, dtype=np.float64)
    out, off = [], 0
    for i, o in LAYERS:
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W, b))
    assert off == N_PARAMS
    return out


def forward(p, X):
    """Outputs (B,) and the activations [X, H1, H2, H3] (post-ReLU)."""
    acts = [np.asarray(X, dtype=np.float64)]
    layers = unpack(p)
    h = acts[0]
    for li, (W, b) in enumerate(layers):
        z = h @ W.T + b
        if li < len(layers) - 1:
            h = np.maximum(z, 0.0)
            acts.append(h)
        else:
            h = z[:, 0]
    return h, acts


def loss(p, X, y):
    out, _ = forward(p, X)
    return float(np.mean((out - np.asarray(y, dtype=np.float64)) ** 2))


def loss_and_grad(p, X, y):
    """Mean squared error and its gradient (flat, same layout as p)."""
    layers = unpack(p)
    out, acts = forward(p, X)
    B = len(out)
    d = (2.0 / B) * (out - np.asarray(y, dtype=np.float64))
    dz = d[:, None]  # gradient w.r.t. the last layer's pre-activation
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        h_in = acts[li]
        grads[li] = (dz.T @ h_in, dz.sum(axis=0))
        if li > 0:
            dz = (dz @ W) * (h_in > 0.0)  # ReLU'(z) = [z > 0] = [h > 0]
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return float(np.mean((out - y) ** 2)), flat


def finite_difference_grad(p, X, y, idx, h=1e-6):
    """Central differences of the loss at the parameters idx."""
    p = np.asarray(p, dtype=np.float64).copy()
    g = np.empty(len(idx))
    for n, k in enumerate(idx):
        v = p[k]
        p[k] = v + h
        lp = loss(p, X, y)
        p[k] = v - h
        lm = loss(p, X, y)
        p[k] = v
        g[n] = (lp - lm) / (2 * h)
    return g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) at step t >= 1;
    returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed):
    """He-uniform weights and zero biases (CPU tests only; the library
    draws its own)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in LAYERS:
        lim = np.sqrt(6.0 / i)
        parts.append(rng.uniform(-lim, lim, size=o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts)
