"""The cost-model oracle (oracle/mlp.py) pinned by central finite differences, and the tile
policy API's argument checks (no GPU call)."""
import numpy as np
import pytest

from oracle import mlp as OM


def _data(n=40, din=7, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, din))
    t = np.sin(X[:, 0]) + 0.3 * X[:, 1] ** 2
    return X, t


@pytest.mark.parametrize("din,hidden", [(7, 5), (12, 16)])
def test_oracle_gradient_finite_differences(din, hidden):
    X, t = _data(din=din)
    rng = np.random.default_rng(1)
    p = rng.standard_normal(OM.num_params(din, hidden)) * 0.5
    loss, g = OM.loss_and_grad(p, X, t, din, hidden)
    h = 1e-6
    for i in rng.choice(len(p), size=min(len(p), 60), replace=False):
        e = np.zeros_like(p)
        e[i] = h
        fd = (OM.loss_and_grad(p + e, X, t, din, hidden)[0] - OM.loss_and_grad(p - e, X, t, din, hidden)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd))


def test_oracle_forward_closed_form():
    """One hidden unit per layer, identity-like weights: the output is relu(relu(x)) + bias."""
    din, hidden = 1, 1
    p = np.array([1.0, 0.0, 1.0, 0.0, 2.0, 0.5])   # W1 b1 W2 b2 W3 b3
    X = np.array([[-1.0], [0.0], [3.0]])
    y, _ = OM.forward(p, X, din, hidden)
    assert np.allclose(y, [0.5, 0.5, 6.5])


def test_oracle_adam_decreases_loss():
    X, t = _data(din=6)
    p0 = np.random.default_rng(2).standard_normal(OM.num_params(6, 8)) * 0.3
    l0 = OM.loss_and_grad(p0, X, t, 6, 8)[0]
    _, l1 = OM.adam(p0, X, t, 6, 8, epochs=200, lr=1e-2)
    assert l1 < 0.5 * l0


def test_tile_policy_argument_checks():
    from pureumfpack_hip import _lib
    lib = _lib.load()
    nt = np.zeros(1, np.int32); nb = np.zeros(1, np.int32)
    tm = np.zeros(16, np.int32); tn = np.zeros(16, np.int32); hi = np.zeros(16, np.int32)
    assert lib.mf_gemm_tile_info(nt.ctypes.data_as(_lib.c_i32p), tm.ctypes.data_as(_lib.c_i32p),
                                 tn.ctypes.data_as(_lib.c_i32p), nb.ctypes.data_as(_lib.c_i32p),
                                 hi.ctypes.data_as(_lib.c_i32p)) == 0
    assert nt[0] >= 2 and tm[0] == 64 and tn[0] == 64
    bad = np.full(2 * nb[0], nt[0], np.int32)       # tile index out of range
    assert lib.mf_set_gemm_tiles(1, bad.ctypes.data_as(_lib.c_i32p)) != 0
    ok = np.zeros(2 * nb[0], np.int32)
    assert lib.mf_set_gemm_tiles(1, ok.ctypes.data_as(_lib.c_i32p)) == 0
    got = np.ones(2 * nb[0], np.int32)
    assert lib.mf_get_gemm_tiles(1, got.ctypes.data_as(_lib.c_i32p)) == 0
    assert np.array_equal(got, ok)


def test_oracle_adam_closed_form_steps():
    """Adam's first step with bias correction is p - lr g / (|g| + eps) for every parameter
    (m_hat = g, v_hat = g^2).  The second step, from the two gradients g1 and g2, uses
    m_hat = (0.09 g1 + 0.1 g2) / (1 - 0.9^2) and v_hat = (0.000999 g1^2 + 0.001 g2^2) /
    (1 - 0.999^2).  An Adam without the bias correction moves the first step by
    0.1 lr / sqrt(0.001), about 3.2 lr, and fails."""
    X, t = _data(din=5)
    p0 = np.random.default_rng(4).standard_normal(OM.num_params(5, 6)) * 0.4
    lr, eps = 1e-3, 1e-8
    _, g1 = OM.loss_and_grad(p0, X, t, 5, 6)
    p1, _ = OM.adam(p0, X, t, 5, 6, epochs=1, lr=lr)
    assert np.allclose(p1, p0 - lr * g1 / (np.abs(g1) + eps), rtol=0, atol=1e-15)
    _, g2 = OM.loss_and_grad(p1, X, t, 5, 6)
    p2, _ = OM.adam(p0, X, t, 5, 6, epochs=2, lr=lr)
    m2 = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.9 ** 2)
    v2 = (0.999 * 0.001 * g1 * g1 + 0.001 * g2 * g2) / (1 - 0.999 ** 2)
    assert np.allclose(p2, p1 - lr * m2 / (np.sqrt(v2) + eps), rtol=0, atol=1e-14)
