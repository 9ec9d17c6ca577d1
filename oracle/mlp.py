"""CPU oracle of the GPU cost-model network (test infrastructure: only tests/ use it).

Restates, in float64 NumPy, what pureumfpack.jl_amd/csrc/mf_mlp.hip computes: a
d_in -> hidden (ReLU) -> hidden (ReLU) -> 1 network with parameters laid out per layer as
W (d_out x d_in, row-major) followed by b, the mean-squared-error loss, its gradient by
backpropagation, and full-batch Adam (Kingma & Ba 2015, Algorithm 1).  The reference has no
such component (its tile shapes are fixed); this oracle is pinned by central finite differences
of its own loss (tests/test_oracle_mlp.py), not by reference data.
"""
import numpy as np


def layer_dims(din, hidden):
    return [(din, hidden), (hidden, hidden), (hidden, 1)]


def num_params(din, hidden):
    return sum(i * o + o for i, o in layer_dims(din, hidden))


def unpack(params, din, hidden):
    """[(W, b)] views of a flat parameter vector."""
    out, off = [], 0
    for i, o in layer_dims(din, hidden):
        W = params[off:off + i * o].reshape(o, i)
        off += i * o
        b = params[off:off + o]
        off += o
        out.append((W, b))
    return out


def forward(params, X, din, hidden):
    """Output (n,) and the per-layer (Z, Y) pairs."""
    p = np.asarray(params, np.float64)
    A = np.asarray(X, np.float64)
    cache = []
    layers = unpack(p, din, hidden)
    for l, (W, b) in enumerate(layers):
        Z = A @ W.T + b
        Y = np.maximum(Z, 0.0) if l < 2 else Z
        cache.append((A, Z))
        A = Y
    return A[:, 0], cache


def loss_and_grad(params, X, t, din, hidden):
    """Mean squared error and its gradient with respect to the flat parameters."""
    p = np.asarray(params, np.float64)
    y, cache = forward(p, X, din, hidden)
    t = np.asarray(t, np.float64)
    n = len(t)
    d = y - t
    loss = float(d @ d / n)
    dZ = (2.0 / n) * d[:, None]
    grads = []
    layers = unpack(p, din, hidden)
    for l in range(2, -1, -1):
        A, Z = cache[l]
        if l < 2:
            dZ = dZ * (Z > 0)
        W, _ = layers[l]
        grads.append((dZ.T @ A, dZ.sum(axis=0)))
        dZ = dZ @ W
    g = []
    for dW, db in reversed(grads):
        g.append(dW.ravel())
        g.append(db)
    return loss, np.concatenate(g)


def adam(params, X, t, din, hidden, epochs, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Full-batch Adam from the given parameters; returns (parameters, final loss)."""
    p = np.array(params, np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, epochs + 1):
        _, g = loss_and_grad(p, X, t, din, hidden)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** step)
        vh = v / (1 - b2 ** step)
        p -= lr * mh / (np.sqrt(vh) + eps)
    return p, loss_and_grad(p, X, t, din, hidden)[0]
