"""The learned tile cost model on the GPU: its forward/backward/Adam kernels against the
oracle, tile shapes leave results bit-identical, and what the trained network gains."""
import numpy as np
import pytest

from oracle import mlp as OM
from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from pureumfpack_hip import tuning as T

pytestmark = pytest.mark.gpu


def _data(n, din, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, din)).astype(np.float32)
    t = (np.sin(X[:, 0]) + 0.3 * X[:, 1] ** 2 - 0.2 * X[:, 2]).astype(np.float32)
    return X, t


@pytest.mark.parametrize("n,din,hidden", [(64, 12, 16), (300, 12, 32)])
def test_mlp_forward_and_gradient_match_oracle(n, din, hidden):
    X, t = _data(n, din)
    net = T.MLP(din, hidden, seed=3)
    p = net.get_params().astype(np.float64)
    y = net.forward(X)
    yo, _ = OM.forward(p, X, din, hidden)
    assert np.abs(y - yo).max() <= 1e-5 * max(1.0, np.abs(yo).max())
    loss, g = net.gradient(X, t)
    lo, go = OM.loss_and_grad(p, X, t, din, hidden)
    assert abs(loss - lo) <= 1e-5 * max(1.0, lo)
    assert np.abs(g - go).max() <= 1e-4 * max(1e-3, np.abs(go).max())


def test_mlp_adam_matches_oracle():
    din, hidden = 12, 16
    X, t = _data(128, din, seed=1)
    net = T.MLP(din, hidden, seed=5)
    p0 = net.get_params().astype(np.float64)
    loss0 = net.gradient(X, t)[0]
    loss = net.train(X, t, epochs=60, lr=5e-3)
    po, lo = OM.adam(p0, X, t, din, hidden, epochs=60, lr=5e-3)
    assert loss < 0.7 * loss0
    assert abs(loss - lo) <= 1e-3 * max(1.0, lo)
    assert np.abs(net.get_params() - po).max() <= 1e-3


def test_tile_policy_leaves_factors_bit_identical():
    A = M.poisson3d(16)
    S = I.symbolic_mf(A)
    ref = None
    saved = T.get_policy("fp64")
    try:
        for tile in range(len(T.tile_info()[0])):
            T.set_policy("fp64", np.full_like(saved, tile))
            F = I.splu(A, method="multifrontal", symbolic=S, small_nf=1)
            cur = (F.L.data.copy(), F.U.data.copy())
            if ref is None:
                ref = cur
            assert np.array_equal(cur[0], ref[0]) and np.array_equal(cur[1], ref[1])
    finally:
        T.set_policy("fp64", saved)


def test_cost_model_gain_over_measured_autotuning():
    """The model-driven tuner reaches (nearly) the exhaustive policy's speed with fewer timings,
    and never a policy slower than the default tiles (measured timing decides)."""
    tu = T.Autotuner("fp32", reps=8)
    pol_m, model = tu.with_model(frac=0.4, seed=0)
    n_model = tu.timings
    tu2 = T.Autotuner("fp32", reps=8)
    pol_x = tu2.exhaustive()
    n_exh = tu2.timings
    default = np.zeros_like(pol_x)
    t_def = tu.policy_time(default, reps=20)
    t_mod = tu.policy_time(pol_m, reps=20)
    t_exh = tu.policy_time(pol_x, reps=20)
    print(f"[tuning] timings model {n_model} vs exhaustive {n_exh}; total s default {t_def:.3e} "
          f"model {t_mod:.3e} exhaustive {t_exh:.3e}; policy model {pol_m.tolist()} exh {pol_x.tolist()}",
          flush=True)
    assert n_model < n_exh
    assert t_mod <= 1.05 * t_def
    assert t_mod <= 1.15 * t_exh


def test_plan_tuner_on_the_real_schur_launches():
    """Tuning on the plan's own Schur-complement launches (3D Poisson 32^3, every level's gemm
    problems by size bucket): the model-driven policy needs fewer timings than the exhaustive
    one and is never slower than the default tiles or (beyond timing noise) the exhaustive
    pick; measurement overrules every wrong prediction."""
    S = I.symbolic_mf(M.poisson3d(32))
    tx = T.PlanTuner(S, "fp32")
    pol_x = tx.exhaustive()
    tm = T.PlanTuner(S, "fp32")
    pol_m = tm.with_model(seed=1)
    default = np.zeros_like(pol_x)
    t_def, t_exh, t_mod = (tx.policy_time(p) for p in (default, pol_x, pol_m))
    print(f"[tuning] plan launches: timings model {tm.timings} vs exhaustive {tx.timings}; Schur ms default "
          f"{t_def * 1e3:.3f} exhaustive {t_exh * 1e3:.3f} model {t_mod * 1e3:.3f}; policies {pol_x[1].tolist()} "
          f"{pol_m[1].tolist()}", flush=True)
    assert tm.timings < tx.timings
    assert t_exh <= t_def
    assert t_mod <= 1.02 * t_def and t_mod <= 1.05 * t_exh
