"""CPU oracle of the learned closure (mercurydpm_fork_amd/csrc/closure_mlp.hip).
TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the HIP kernels,
never by the product.  A numpy FP64 restatement of the network 4 -> 32 -> 32
-> 2 (tanh hidden layers, linear output), of the mean-squared-error loss
(1/n) sum_s sum_k (y_sk - t_sk)^2 and of its gradient by back-propagation.
The closure has no counterpart in the reference; the gradient is pinned by
central finite differences (tests/test_closure.py)."""
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
