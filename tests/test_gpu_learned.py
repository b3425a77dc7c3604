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


def test_inference_kernel_equals_network_on_features():
    """the solver's inference kernel (BF16 matrix cores) equals the CPU reference of its arithmetic
    (features and first-layer activations rounded to BF16, FP32 accumulation); a different summation
    order can move a first-layer activation across a BF16 rounding boundary, so a few cells may
    differ by more, and every cell stays within the BF16 rounding of the FP32 network"""
    from oracle import mlp_ref
    from plens_amd import GpuPlens, shock_tube
    from plens_amd.learned import MLP, init_params, samples, unpack, use_network
    case = shock_tube(nx=16, ny=2, nz=2)
    g = GpuPlens(case)
    X, a_plain = samples(g)
    m = MLP(init_params(7))
    floor = 0.02

    def clip(y):
        y = y.astype(np.float64)
        y[y < floor] = 0.0
        return np.minimum(y, case.physics.alpha_max)
    y = clip(mlp_ref.infer(unpack(m.params()), X))
    y32 = clip(m.predict(X))  # the FP32 network of the training kernels
    use_network(g, m, tolerance=1.0, floor=floor)  # never fall back
    X2, a_net = samples(g)
    assert np.array_equal(X, X2)
    d = np.abs(a_net - y)
    print(f"inference vs BF16 reference: max {d.max():.2e}, 99th percentile {np.quantile(d, 0.99):.2e}; "
          f"vs FP32 network max {np.abs(a_net - y32).max():.2e}", flush=True)
    assert np.quantile(d, 0.99) < 1e-6 and d.max() < 2e-3
    assert np.abs(a_net - y32).max() < 1e-2
    use_network(g, m, tolerance=0.05, floor=floor)
    _, a_mix = samples(g)
    fb = np.abs(a_net - a_plain) > 0.05  # the fallback compares the kernel's own value
    assert np.array_equal(a_mix[fb], a_plain[fb]) and np.array_equal(a_mix[~fb], a_net[~fb])
    use_network(g, None)
    _, a_back = samples(g)
    assert np.array_equal(a_back, a_plain)


def test_training_reduces_loss_on_solver_samples():
    from plens_amd import shock_tube
    from plens_amd.learned import MLP, collect
    X, y = collect(shock_tube(nx=32, ny=2, nz=2), steps=40, every=10)
    m = MLP(seed=0)
    hist = m.train(X, y, epochs=30, batch=64, lr=1e-3, seed=1)
    print("loss", hist[0], hist[-1], flush=True)
    assert hist[-1] < 0.5 * hist[0]


def test_trained_network_gains_over_untrained():
    """what training buys: on a held-out shock-tube seed the trained network agrees with the plain
    blender far better than the untrained one (lower error, fewer fallbacks), and the learned run
    stays close to the plain run"""
    from plens_amd import GpuPlens, shock_tube
    from plens_amd.indicator_study import agreement, shock_position
    from plens_amd.learned import MLP, collect, init_params, stats, use_network
    Xs, ys = [], []
    for s in range(1, 5):
        X, y = collect(shock_tube(nx=32, ny=2, nz=2, seed=s), steps=60, every=10)
        Xs.append(X)
        ys.append(y)
    X, y = np.concatenate(Xs), np.concatenate(ys)
    held = shock_tube(nx=32, ny=2, nz=2, seed=1001)
    Xh, yh = collect(held, steps=60, every=10)
    untrained = MLP(init_params(0))
    trained = MLP(init_params(0))
    trained.train(X, y, epochs=300, batch=256, lr=1e-3, seed=2)
    a_u, a_t = agreement(untrained, Xh, yh), agreement(trained, Xh, yh)
    print(f"untrained {a_u}\ntrained {a_t}", flush=True)
    assert a_t["mae"] < 0.2 * a_u["mae"]
    assert a_t["fallback_fraction"] < 0.1 * a_u["fallback_fraction"]
    plain, learned = GpuPlens(held), GpuPlens(held)
    use_network(learned, trained)
    plain.run(60, held.courant)
    learned.run(60, held.courant)
    ev, fb = stats(learned)
    Qp, Ql = plain.get_state(), learned.get_state()
    drho = np.linalg.norm(Ql[0] - Qp[0]) / np.linalg.norm(Qp[0])
    dx = 4.0 / 32
    ds = abs(shock_position(held, Ql) - shock_position(held, Qp)) / dx
    print(f"learned vs plain: density rel L2 {drho:.3e}, shock position {ds:.4f} cells, "
          f"fallbacks {fb}/{ev}", flush=True)
    assert drho < 1e-2 and ds < 0.1


def test_trained_network_blends_fewer_cells_than_plain_indicator():
    """the gain over the plain method: on the held-out shock tube (seed 1001, 128x8x8, N=3, its 200
    steps) the trained indicator (network of plens_amd/indicator_mlp.npy, plain fallback at 0.05)
    sends fewer (cell, stage) pairs through the first-order subcell blend than the plain indicator
    while the shock stays where the plain run puts it"""
    from plens_amd import GpuPlens, shock_tube
    from plens_amd.indicator_study import HELD_OUT_SEED, load_trained, shock_position
    from plens_amd.learned import use_network
    case = shock_tube(seed=HELD_OUT_SEED)
    plain, learned = GpuPlens(case), GpuPlens(case)
    use_network(learned, load_trained(), 0.05)
    plain.run(200, case.courant)
    learned.run(200, case.courant)
    bp, bl = plain.blended_cell_stages(), learned.blended_cell_stages()
    Qp, Ql = plain.get_state(), learned.get_state()
    ds = abs(shock_position(case, Ql) - shock_position(case, Qp)) / (4.0 / 128)
    drho = np.linalg.norm(Ql[0] - Qp[0]) / np.linalg.norm(Qp[0])
    print(f"blended cell-stages plain {bp}, learned {bl} ({bl / bp:.3f}); shock position difference "
          f"{ds:.2e} cells, density rel L2 {drho:.2e}", flush=True)
    assert bl < 0.9 * bp
    assert ds < 0.01 and drho < 2e-3
