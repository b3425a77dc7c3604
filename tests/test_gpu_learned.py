"""GPU checks of the learned troubled-cell indicator: the hand-written forward / backward / Adam
kernels against the CPU reference network (oracle/mlp_ref.py), the fused inference in the solver's
blender kernel, and a short training run."""
import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _ref_params(P):
    from plens_amd.learned import unpack
    return unpack(P)


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, 125)) * 0.2).astype(np.float32)
    X[:, 0] = 1.0
    t = rng.uniform(0, 0.5, n).astype(np.float32)
    return X, t


def test_forward_and_gradients_match_reference():
    from plens_amd.learned import MLP, init_params
    P = init_params(3)
    P[-65:-1] *= 10  # larger output weights so every gradient block is well above rounding
    m = MLP(P)
    X, t = _batch(512, 4)
    ref = _ref_params(P)
    y_ref, _ = M.forward(ref, X)
    y = m.predict(X)
    assert np.abs(y - y_ref).max() < 2e-6
    loss, g = m.grad(X, t)
    loss_ref, g_ref = M.loss_and_grads(ref, X, t)
    assert abs(loss - loss_ref) < 1e-6 * max(1.0, loss_ref)
    flat = np.concatenate([g_ref[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "w3", "b3")])
    err = np.abs(g - flat).max() / np.abs(flat).max()
    print(f"grad rel err {err:.2e}", flush=True)
    assert err < 1e-4


def test_adam_step_matches_reference():
    from plens_amd.learned import MLP, init_params
    P = init_params(5)
    m = MLP(P)
    X, t = _batch(256, 6)
    m.step(X, t, lr=1e-3)
    ref = _ref_params(P)
    _, g = M.loss_and_grads(ref, X, t)
    new = M.adam_step(ref, g, {}, lr=1e-3)
    flat = np.concatenate([new[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "w3", "b3")])
    got = m.params()
    moved = np.abs(flat - P) > 5e-4  # the signed first step moved these by ~lr
    assert moved.mean() > 0.5
    assert np.abs(got - flat).max() < 2e-5


def test_fused_inference_equals_network_on_features():
    from plens_amd import GpuPlens, shock_tube
    from plens_amd.learned import MLP, init_params, samples, use_network
    case = shock_tube(nx=16, ny=2, nz=2)
    g = GpuPlens(case)
    X, a_plain = samples(g)
    m = MLP(init_params(7))
    y = m.predict(X).astype(np.float64)
    y[y < case.physics.alpha_min] = 0.0
    use_network(g, m, tolerance=1.0)  # never fall back
    X2, a_net = samples(g)
    assert np.array_equal(X, X2)
    assert np.abs(a_net - y).max() < 1e-6
    use_network(g, m, tolerance=0.05)
    _, a_mix = samples(g)
    fb = np.abs(y - a_plain) > 0.05
    assert np.allclose(a_mix[fb], a_plain[fb]) and np.allclose(a_mix[~fb], y[~fb], atol=1e-6)


def test_training_reduces_loss_on_solver_samples():
    from plens_amd import shock_tube
    from plens_amd.learned import MLP, collect
    X, y = collect(shock_tube(nx=32, ny=2, nz=2), steps=40, every=10)
    m = MLP(seed=0)
    hist = m.train(X, y, epochs=30, batch=64, lr=1e-3, seed=1)
    print("loss", hist[0], hist[-1], flush=True)
    assert hist[-1] < 0.5 * hist[0]
