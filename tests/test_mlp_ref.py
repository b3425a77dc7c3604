"""Pins of the CPU reference of the learned-indicator network (oracle/mlp_ref.py)."""
import numpy as np

from oracle import mlp_ref as M


def test_bf16_rounding_known_values():
    # bf16 spacing at 1 is 2^-7; halfway cases go to the even mantissa
    x = np.array([1.0, 1 + 2 ** -8, 1 + 2 ** -7 + 2 ** -8, 1 + 2 ** -8 + 2 ** -10, -2.5, 3.1415926], np.float32)
    got = M.bf16(x)
    expect = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 3.140625], np.float32)
    assert np.array_equal(got, expect)


def test_backprop_matches_finite_differences(monkeypatch):
    """analytic gradients of the MSE loss against central differences, in float64 with the BF16
    rounding switched off (so the check sees only the backpropagation formulas)"""
    monkeypatch.setattr(M, "bf16", lambda x: np.asarray(x, np.float64))
    rng = np.random.default_rng(0)
    p = {k: v.astype(np.float64) for k, v in M.init_params(1).items()}
    p["w3"] *= 10
    X = rng.standard_normal((16, M.IN)) * 0.3
    t = rng.uniform(0, 0.5, 16)
    old = (M.GELU_C, M.GELU_A)
    M.GELU_C, M.GELU_A = np.float64(old[0]), np.float64(old[1])
    try:
        def loss(pp):
            y, _ = M.forward(pp, X.astype(np.float64))
            return np.mean((y - t) ** 2)
        M_forward = M.forward
        def fwd64(pp, XX):  # forward without the float32 cast of the inputs
            z1 = XX @ pp["W1"].T + pp["b1"]
            a1 = M.gelu(z1)
            z2 = a1 @ pp["W2"].T + pp["b2"]
            a2 = M.gelu(z2)
            z3 = a2 @ pp["w3"] + pp["b3"][0]
            y = 0.5 / (1 + np.exp(-z3))
            return y, (XX, z1, a1, z2, a2, z3, y)
        monkeypatch.setattr(M, "forward", fwd64)
        _, g = M.loss_and_grads(p, X, t)
        for k in p:
            flat = p[k].reshape(-1)
            for i in np.random.default_rng(2).choice(flat.size, size=min(6, flat.size), replace=False):
                h = 1e-6
                old_v = flat[i]
                flat[i] = old_v + h
                lp = np.mean((fwd64(p, X)[0] - t) ** 2)
                flat[i] = old_v - h
                lm = np.mean((fwd64(p, X)[0] - t) ** 2)
                flat[i] = old_v
                fd = (lp - lm) / (2 * h)
                an = g[k].reshape(-1)[i]
                assert abs(fd - an) <= 1e-6 * max(1e-3, abs(fd)), (k, i, fd, an)
        assert M_forward is not None
    finally:
        M.GELU_C, M.GELU_A = old


def test_first_adam_step_is_signed_lr():
    p = {"a": np.array([0.5, -0.25, 1.0], np.float32)}
    g = {"a": np.array([3.0, -0.01, 0.0], np.float32)}
    st = {}
    q = M.adam_step({k: v.copy() for k, v in p.items()}, g, st, lr=1e-3)
    # bias-corrected first step: m_hat = g, v_hat = g^2 -> lr * g / (|g| + eps)
    expect = p["a"] - 1e-3 * g["a"] / (np.abs(g["a"]) + 1e-8)
    assert np.allclose(q["a"], expect, rtol=0, atol=1e-7)


def test_inference_arithmetic_rounds_the_matrix_operands_only():
    """infer() is forward() with BF16 operands of the two hidden-layer products: rounding the
    features first changes nothing, the result moves from the FP32 network by no more than BF16
    rounding of those operands allows, and it does move (the rounding is applied)"""
    rng = np.random.default_rng(3)
    p = M.init_params(5)
    p["b1"] = rng.normal(0, 0.1, p["b1"].shape).astype(np.float32)
    X = rng.normal(0, 0.3, (256, M.IN)).astype(np.float32)
    y = M.infer(p, X)
    assert np.array_equal(y, M.infer(p, M.bf16(X)))
    y32, _ = M.forward(p, X)
    d = np.abs(y - y32)
    assert d.max() < 5e-3 and d.max() > 0
    # with a zero first layer the first-layer activations are gelu(b1) for every cell: layer 2 then
    # sees exactly their BF16 rounding
    Xb = M.bf16(X)
    p0 = dict(p)
    p0["W1"] = np.zeros_like(p["W1"])
    a1 = M.gelu(p0["b1"])
    y0 = M.infer(p0, Xb)
    z2 = M.bf16(a1)[None, :] @ M.bf16(p0["W2"]).T + p0["b2"]
    z3 = M.gelu(z2) @ M.bf16(p0["w3"]) + p0["b3"][0]
    ref = np.float32(0.5) / (np.float32(1) + np.exp(-z3))
    assert np.allclose(y0, ref[0], rtol=0, atol=1e-7)
