"""CPU oracle of the learned closure (mercurydpm_fork_amd/csrc/closure_mlp.hip
and closure_mfma.hip).  TEST INFRASTRUCTURE ONLY: used by tests/ as the
checker of the HIP kernels, never by the product.  numpy restatements of the
networks 4 -> 32 -> 32 -> 2 (tanh, FP64) and 4 -> 3 x 128 -> 2 (GELU, BF16
weights), of the mean-squared-error loss (1/n) sum_s sum_k (y_sk - t_sk)^2
and of its gradient by back-propagation.  The closure has no counterpart in
the reference; the gradients are pinned by central finite differences and the
BF16 rounding by exact values (tests/test_closure.py)."""
import numpy as np

H = 32
SHAPES = [("W1", (H, 4)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (2, H)), ("b3", (2,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[k:k + n].reshape(shape)
        k += n
    return out


def forward(p, x):
    w = unpack(p)
    h1 = np.tanh(x @ w["W1"].T + w["b1"])
    h2 = np.tanh(h1 @ w["W2"].T + w["b2"])
    return h2 @ w["W3"].T + w["b3"], h1, h2


def loss_and_gradient(p, x, t):
    w = unpack(p)
    n = len(x)
    y, h1, h2 = forward(p, x)
    e = y - t
    loss = float(np.sum(e * e) / n)
    d3 = 2.0 * e / n
    d2 = (d3 @ w["W3"]) * (1.0 - h2 * h2)
    d1 = (d2 @ w["W2"]) * (1.0 - h1 * h1)
    g = [d1.T @ x, d1.sum(0), d2.T @ h1, d2.sum(0), d3.T @ h2, d3.sum(0)]
    return loss, np.concatenate([a.ravel() for a in g])


def init(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    return np.concatenate(parts)


# ---------------------------------------------------------------------------
# The configured network (mercurydpm_fork_amd/csrc/closure_mfma.hip):
# 4 -> 128 -> 128 -> 128 -> 2, GELU hidden layers, BF16 weights and BF16
# activations into the 128 x 128 layers, FP32 elsewhere.  forward128 /
# loss_and_gradient128 with quantize=True restate the kernels' arithmetic in
# numpy float32 (sums in another order: the comparison allows for FP32
# rounding); with quantize=False, in float64 without rounding, they are the
# exact network whose gradient the finite differences pin.
# ---------------------------------------------------------------------------
from math import sqrt as _sqrt

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(__import__("math").erf)

H128 = 128
SHAPES128 = [("W1", (H128, 4)), ("b1", (H128,)), ("W2", (H128, H128)), ("b2", (H128,)), ("W3", (H128, H128)),
             ("b3", (H128,)), ("W4", (2, H128)), ("b4", (2,))]
N_PARAMS128 = sum(int(np.prod(s)) for _, s in SHAPES128)


def unpack128(p):
    out, k = {}, 0
    for name, shape in SHAPES128:
        n = int(np.prod(shape))
        out[name] = p[k:k + n].reshape(shape)
        k += n
    return out


def bf16(a):
    """float32 -> bfloat16 -> float32, round to nearest even (finite values)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / _sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / _sqrt(2.0))) + z * np.exp(-0.5 * z * z) / _sqrt(2.0 * np.pi)


def forward128(p, x, quantize=True):
    dt = np.float32 if quantize else np.float64
    q = bf16 if quantize else (lambda a: a)
    w = {k: np.asarray(v, dt) for k, v in unpack128(np.asarray(p, dt)).items()}
    x = np.asarray(x, dt)
    wb = {k: (q(v) if k.startswith("W") else v) for k, v in w.items()}
    z1 = x @ wb["W1"].T + w["b1"]
    h1 = q(gelu(z1).astype(dt))
    z2 = h1 @ wb["W2"].T + w["b2"]
    h2 = q(gelu(z2).astype(dt))
    z3 = h2 @ wb["W3"].T + w["b3"]
    h3 = q(gelu(z3).astype(dt))
    y = h3 @ wb["W4"].T + w["b4"]
    return y, (x, z1, h1, z2, h2, z3, h3, wb)


def loss_and_gradient128(p, x, t, quantize=True):
    dt = np.float32 if quantize else np.float64
    q = bf16 if quantize else (lambda a: a)
    y, (x, z1, h1, z2, h2, z3, h3, wb) = forward128(p, x, quantize)
    n = len(x)
    e = y - np.asarray(t, dt)
    loss = float(np.sum(e * e) / n)
    dy = (2.0 * e / n).astype(dt)
    d3 = (dy @ wb["W4"]) * gelu_grad(z3).astype(dt)
    e3 = q(d3.astype(dt))
    d2 = (e3 @ wb["W3"]) * gelu_grad(z2).astype(dt)
    e2 = q(d2.astype(dt))
    d1 = (e2 @ wb["W2"]) * gelu_grad(z1).astype(dt)
    g = [d1.T @ x, d1.sum(0), e2.T @ h1, d2.sum(0), e3.T @ h2, d3.sum(0), dy.T @ h3, dy.sum(0)]
    return loss, np.concatenate([np.asarray(a, np.float64).ravel() for a in g])


def init128(seed=0):
    """Glorot-uniform weights, zero biases (float32 values)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES128:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    return np.concatenate(parts).astype(np.float32)
