"""TEST INFRASTRUCTURE ONLY: CPU reference of the learned troubled-cell indicator network.

MLP 125-64-64-1 with GELU (tanh form) hidden activations and a 0.5*sigmoid output, weights rounded to
BF16 (round-to-nearest-even) for the forward pass, FP32 accumulation, FP32 master weights and Adam
(lr 1e-3, beta1 0.9, beta2 0.999, eps 1e-8), mean-squared-error loss against the plain blender value.
This is the checker of the GPU forward / backward / Adam kernels (plens_amd/csrc/mlp.hip).
"""
import numpy as np

IN, H1, H2 = 125, 64, 64
GELU_C = np.float32(0.7978845608028654)  # sqrt(2/pi)
GELU_A = np.float32(0.044715)


def bf16(x):
    """round float32 values to bfloat16 (nearest even) and back to float32"""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    r = ((b + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(x):
    t = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return np.float32(0.5) * x * (np.float32(1) + t)


def gelu_grad(x):
    u = GELU_C * (x + GELU_A * x ** 3)
    t = np.tanh(u)
    du = GELU_C * (np.float32(1) + np.float32(3) * GELU_A * x * x)
    return np.float32(0.5) * (np.float32(1) + t) + np.float32(0.5) * x * (np.float32(1) - t * t) * du


def init_params(seed):
    """He-style uniform init in FP32 (the GPU kernels receive these exact arrays)."""
    rng = np.random.default_rng(seed)
    def u(fan_in, shape):
        lim = np.float32(np.sqrt(6.0 / fan_in))
        return rng.uniform(-lim, lim, shape).astype(np.float32)
    return {"W1": u(IN, (H1, IN)), "b1": np.zeros(H1, np.float32),
            "W2": u(H1, (H2, H1)), "b2": np.zeros(H2, np.float32),
            "w3": u(H2, (H2,)) * np.float32(0.1), "b3": np.zeros(1, np.float32)}


def forward(p, X):
    X = X.astype(np.float32)
    z1 = X @ bf16(p["W1"]).T + p["b1"]
    a1 = gelu(z1)
    z2 = a1 @ bf16(p["W2"]).T + p["b2"]
    a2 = gelu(z2)
    z3 = a2 @ bf16(p["w3"]) + p["b3"][0]
    y = np.float32(0.5) / (np.float32(1) + np.exp(-z3))
    return y, (X, z1, a1, z2, a2, z3, y)


def infer(p, X):
    """the solver-side inference arithmetic: as forward(), with the features and the first hidden
    layer's activations rounded to BF16 before they enter the matrix products (the operands of BF16
    matrix cores), FP32 accumulation; the output layer in FP32"""
    Xb = bf16(X.astype(np.float32))
    a1 = bf16(gelu(Xb @ bf16(p["W1"]).T + p["b1"]))
    a2 = gelu(a1 @ bf16(p["W2"]).T + p["b2"])
    z3 = a2 @ bf16(p["w3"]) + p["b3"][0]
    return np.float32(0.5) / (np.float32(1) + np.exp(-z3))


def loss_and_grads(p, X, t):
    y, (X, z1, a1, z2, a2, z3, y) = forward(p, X)
    n = np.float32(X.shape[0])
    r = y - t.astype(np.float32)
    loss = float(np.mean(r * r))
    dy = np.float32(2) * r / n
    dz3 = dy * y * (np.float32(1) - np.float32(2) * y)  # d(0.5 s)/dz = 0.5 s (1-s) = y (1 - 2y)
    g = {"w3": a2.T @ dz3, "b3": np.array([dz3.sum()], np.float32)}
    da2 = np.outer(dz3, bf16(p["w3"]))
    dz2 = da2 * gelu_grad(z2)
    g["W2"] = dz2.T @ a1
    g["b2"] = dz2.sum(0)
    da1 = dz2 @ bf16(p["W2"])
    dz1 = da1 * gelu_grad(z1)
    g["W1"] = dz1.T @ X
    g["b1"] = dz1.sum(0)
    return loss, {k: v.astype(np.float32) for k, v in g.items()}


def adam_step(p, g, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in p:
        m = state.setdefault("m_" + k, np.zeros_like(p[k]))
        v = state.setdefault("v_" + k, np.zeros_like(p[k]))
        m[:] = np.float32(b1) * m + np.float32(1 - b1) * g[k]
        v[:] = np.float32(b2) * v + np.float32(1 - b2) * g[k] * g[k]
        mh = m / np.float32(1 - b1 ** t)
        vh = v / np.float32(1 - b2 ** t)
        p[k] = (p[k] - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32)
    return p
